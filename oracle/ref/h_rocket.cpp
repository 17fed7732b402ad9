/* harness over code/rocket.cpp: the GNU Rocket wrapper in player mode
 * (SYNC_PLAYER), whose tracks load from "sync/_<name>.track" files relative
 * to the working directory (3rdparty/rocket-stripped/lib/device.c). */
#include "code/rocket.cpp"
#include "ref_common.h"
