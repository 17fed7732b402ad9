/* harness over code/rocket.cpp: the GNU Rocket wrapper in player mode
 * (SYNC_PLAYER), whose tracks load from "sync/_<name>.track" files relative
 * to the working directory (3rdparty/rocket-stripped/lib/device.c). */
#include "code/rocket.cpp"
#include "ref_common.h"

#include <unistd.h>

/* Rocket::get of track `name`, read from dir/sync/ (the player started by
 * ref_st_rocket), at `row` */
REF_API double ref_sync_value(const char *dir, const char *name, double row)
{
    char cwd[4096];
    if (!getcwd(cwd, sizeof(cwd)) || chdir(dir) != 0) return -1e300;
    const sync_track *t = Rocket::AddTrack(name);
    if (chdir(cwd) != 0 || !t) return -1e300;
    Rocket::s_rocketRow = row;
    return Rocket::get(t);
}
