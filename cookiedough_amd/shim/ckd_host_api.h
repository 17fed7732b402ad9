// ckd_host_api.h -- what the drop-in shim uses from the host application.
//
// These declarations repeat, with the same names and signatures, the parts
// of the reference's own headers the replaced effect files used:
//   Rocket::AddTrack / Rocket::get   code/rocket.h:8-30
//   Image_Load8 / Image_Load32       code/image.h:11-12
//   PadState / Gamepad_Update        code/gamepad.h:12-22
//   SetLastError                     code/main.h:49
// so the shim links against the application's rocket.cpp, image.cpp,
// gamepad.cpp and main.cpp unchanged.
#pragma once

#include <cstdint>
#include <string>

struct sync_track;
typedef const sync_track *SyncTrack;

namespace Rocket {
const sync_track *AddTrack(const char *name);
double get(const sync_track *track);
} // namespace Rocket

uint32_t *Image_Load32(const std::string &path);
uint8_t *Image_Load8(const std::string &path);

struct PadState {
    float leftX, leftY;
    float rightX, rightY;
    int lShoulder, rShoulder;
};
bool Gamepad_Update(PadState &state);

void SetLastError(const std::string &description);
