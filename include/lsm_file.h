/*
 * lsm_file.h -- compression flags of the reference's file writers
 * (src/utils/lsm_file.h:16-18).  A writer given GZIP or BZIP2 runs
 * `gzip -f` / `bzip2 -f` on the file it wrote; a reader given a name
 * containing ".gz" or ".bz2" decompresses it in place, reads the
 * uncompressed file and compresses it again (src/utils/lsm_file.c:19-70).
 */
#ifndef LSMHIP_LSM_FILE_H
#define LSMHIP_LSM_FILE_H

#define NO_ZIP 0
#define GZIP 1
#define BZIP2 2

#endif
