"""MI355X-native (HIP, gfx950) implementation of LSMLIB's 3D level-set
time-stepping path behind the reference's C entry points."""
from .lib import HipLib, LsmHipError, load  # noqa: F401
