// shim_test.cpp -- drives the C++ drop-in shim (cookiedough_amd/shim) the
// way code/demo.cpp does, with test doubles for the host application's
// Rocket, image loader, gamepad and error reporting:
//   Rocket tracks   values from CKD_SHIM_TRACKS="name=value;name=value"
//   Image_Load8/32  raw files $CKD_SHIM_ASSETS/<file name>.bin
//   Gamepad         always idle
// Usage:
//   shim_test landscape <time> <delta> <out>        (1280x720 frame)
//   shim_test polar <in> <inverse 0|1> <out>
//   shim_test fxblit <in 644x364> <out>
//   shim_test blur07 <in> <W> <H> <strength> <h|v|2d> <inplace 0|1> <out>
//   shim_test tapewarp <in> <W> <H> <strength> <speed> <out>
//   shim_test blend <op> <dst> <src> <n> <arg> <out>       (flat blends, Mix32/Fade32/
//                                                         SoftLight32AA take <arg>)
//   shim_test shadertoy <plasma|nautilus|spikey|spikey_close|sinuses|laura|tunnel>
//             <time> <out>                            (1280x720 frame)
//   shim_test ball <time> <frames> <out>              (frames successive 1280x720
//                                                     frames at the same time)
//   shim_test twister <time> <out>
//   shim_test polar_blit_a <dst> <src> <inverse 0|1> <out>
//   shim_test polar2x2 <in 644x364> <inverse 0|1> <out>
//   shim_test tunnelscape <time> <frames> <out>     (frames successive 1280x720
//                                                     frames at the same time)
//   shim_test testpattern <out>                     (FxBlitter_DrawTestPattern)
//   shim_test lifecycle                             (Create failures and
//             Create/Destroy pairs leave no context reference behind; prints ok)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../cookiedough_amd/shim/ckd_host_api.h"
#include "../../cookiedough_amd/shim/ckd_shim.h"

struct sync_track {
    std::string name;
    double value;
};

static std::map<std::string, sync_track> g_tracks;
static bool g_failed = false;

namespace Rocket {
const sync_track *AddTrack(const char *name)
{
    sync_track &t = g_tracks[name];
    t.name = name;
    t.value = 0.0;
    if (const char *env = getenv("CKD_SHIM_TRACKS")) {
        std::string s = std::string(";") + env, key = std::string(";") + name + "=";
        size_t p = s.find(key);
        if (p != std::string::npos) t.value = atof(s.c_str() + p + key.size());
    }
    return &t;
}
double get(const sync_track *track) { return track ? track->value : 0.0; }
} // namespace Rocket

static std::vector<unsigned char> read_file(const std::string &path)
{
    std::vector<unsigned char> out;
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return out;
    fseek(f, 0, SEEK_END);
    out.resize((size_t)ftell(f));
    fseek(f, 0, SEEK_SET);
    if (fread(out.data(), 1, out.size(), f) != out.size()) out.clear();
    fclose(f);
    return out;
}

static void *load_asset(const std::string &path)
{
    const char *dir = getenv("CKD_SHIM_ASSETS");
    const std::string base = path.substr(path.find_last_of('/') + 1);
    std::vector<unsigned char> d = read_file(std::string(dir ? dir : ".") + "/" + base + ".bin");
    if (d.empty()) return nullptr;
    void *p = malloc(d.size());
    memcpy(p, d.data(), d.size());
    return p;
}

uint32_t *Image_Load32(const std::string &path) { return static_cast<uint32_t *>(load_asset(path)); }
uint8_t *Image_Load8(const std::string &path) { return static_cast<uint8_t *>(load_asset(path)); }

bool Gamepad_Update(PadState &state)
{
    memset(&state, 0, sizeof(state));
    return false;
}

void SetLastError(const std::string &description)
{
    fprintf(stderr, "SetLastError: %s\n", description.c_str());
    g_failed = true;
}

static std::vector<uint32_t> read_pixels(const char *path)
{
    std::vector<unsigned char> d = read_file(path);
    std::vector<uint32_t> px(d.size() / 4);
    memcpy(px.data(), d.data(), px.size() * 4);
    return px;
}

static void write_pixels(const char *path, const uint32_t *px, size_t n)
{
    FILE *f = fopen(path, "wb");
    fwrite(px, 4, n, f);
    fclose(f);
}

int main(int argc, char **argv)
{
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    const size_t n = 1280 * 720;
    std::vector<uint32_t> out(n, 0);
    if (cmd == "landscape" && argc == 5) {
        if (!Landscape_Create()) return 1;
        Landscape_Draw(out.data(), (float)atof(argv[2]), (float)atof(argv[3]));
        Landscape_Destroy();
        write_pixels(argv[4], out.data(), n);
    } else if (cmd == "polar" && argc == 5) {
        std::vector<uint32_t> in = read_pixels(argv[2]);
        if (in.size() != n || !Polar_Create()) return 1;
        Polar_Blit(out.data(), in.data(), atoi(argv[3]) != 0);
        Polar_Destroy();
        write_pixels(argv[4], out.data(), n);
    } else if (cmd == "fxblit" && argc == 4) {
        std::vector<uint32_t> in = read_pixels(argv[2]);
        if (in.size() != 644 * 364 || !FxBlitter_Create()) return 1;
        memcpy(g_pFxMap[0], in.data(), in.size() * 4);
        Fx_Blit_2x2(out.data(), g_pFxMap[0]);
        FxBlitter_Destroy();
        write_pixels(argv[3], out.data(), n);
    } else if (cmd == "blur07" && argc == 9) {
        std::vector<uint32_t> in = read_pixels(argv[2]);
        const unsigned W = (unsigned)atoi(argv[3]), H = (unsigned)atoi(argv[4]);
        const float s = (float)atof(argv[5]);
        const std::string mode = argv[6];
        const bool inplace = atoi(argv[7]) != 0;
        if (in.size() != (size_t)W * H || !BoxBlur_Create()) return 1;
        std::vector<uint32_t> res(in.size());
        uint32_t *dst = inplace ? in.data() : res.data();
        if (mode == "h") HorizontalBoxBlur32(dst, in.data(), W, H, s);
        else if (mode == "v") VerticalBoxBlur32(dst, in.data(), W, H, s);
        else BoxBlur32(dst, in.data(), W, H, s);
        BoxBlur_Destroy();
        write_pixels(argv[8], dst, in.size());
    } else if (cmd == "tapewarp" && argc == 8) {
        std::vector<uint32_t> in = read_pixels(argv[2]);
        const unsigned W = (unsigned)atoi(argv[3]), H = (unsigned)atoi(argv[4]);
        if (in.size() != (size_t)W * H || !BoxBlur_Create()) return 1;
        std::vector<uint32_t> res(in.size());
        TapeWarp32(res.data(), in.data(), W, H, (float)atof(argv[5]), (float)atof(argv[6]));
        BoxBlur_Destroy();
        write_pixels(argv[7], res.data(), res.size());
    } else if (cmd == "blend" && argc == 8) {
        const std::string op = argv[2];
        std::vector<uint32_t> d = read_pixels(argv[3]), s = read_pixels(argv[4]);
        const unsigned n = (unsigned)atoi(argv[5]);
        const float arg = (float)atof(argv[6]);
        if (d.size() < n || s.size() < n || !BoxBlur_Create()) return 1;
        if (op == "mix32") Mix32(d.data(), s.data(), n, (uint8_t)arg);
        else if (op == "mixover32") MixOver32(d.data(), s.data(), n);
        else if (op == "sub32") Sub32(d.data(), s.data(), n);
        else if (op == "excl32") Excl32(d.data(), s.data(), n);
        else if (op == "softlight32") SoftLight32(d.data(), s.data(), n);
        else if (op == "softlight32a") SoftLight32A(d.data(), s.data(), n);
        else if (op == "softlight32aa") SoftLight32AA(d.data(), s.data(), n, arg);
        else if (op == "overlay32") Overlay32(d.data(), s.data(), n);
        else if (op == "overlay32a") Overlay32A(d.data(), s.data(), n);
        else if (op == "darken32_50") Darken32_50(d.data(), s.data(), n);
        else if (op == "mulsrc32") MulSrc32(d.data(), s.data(), n);
        else if (op == "mulsrc32a") MulSrc32A(d.data(), s.data(), n);
        else if (op == "mixsrc32") MixSrc32(d.data(), s.data(), n);
        else if (op == "fade32") Fade32(d.data(), n, s[0], (uint8_t)arg);
        else return 2;
        BoxBlur_Destroy();
        write_pixels(argv[7], d.data(), n);
    } else if (cmd == "shadertoy" && argc == 5) {
        const std::string fx = argv[2];
        const float t = (float)atof(argv[3]);
        if (!Shadertoy_Create()) return 1;
        if (fx == "plasma") Plasma_Draw(out.data(), t, 1.f);
        else if (fx == "nautilus") Nautilus_Draw(out.data(), t, 1.f);
        else if (fx == "spikey") Spikey_Draw(out.data(), t, 1.f, false);
        else if (fx == "spikey_close") Spikey_Draw(out.data(), t, 1.f, true);
        else if (fx == "sinuses") Sinuses_Draw(out.data(), t, 1.f);
        else if (fx == "laura") Laura_Draw(out.data(), t, 1.f);
        else if (fx == "tunnel") Tunnel_Draw(out.data(), t, 1.f);
        else return 2;
        Shadertoy_Destroy();
        write_pixels(argv[4], out.data(), n);
    } else if (cmd == "ball" && argc == 5) {
        const int frames = atoi(argv[3]);
        if (frames < 1 || !Ball_Create()) return 1;
        std::vector<uint32_t> all((size_t)frames * n);
        for (int k = 0; k < frames; ++k) Ball_Draw(all.data() + (size_t)k * n, (float)atof(argv[2]), 1.f);
        if (Ball_GetBackground() == nullptr) return 1;
        printf("has beams: %d\n", Ball_HasBeams() ? 1 : 0);
        Ball_Destroy();
        write_pixels(argv[4], all.data(), all.size());
    } else if (cmd == "twister" && argc == 4) {
        if (!Twister_Create()) return 1;
        Twister_Draw(out.data(), (float)atof(argv[2]), 1.f);
        Twister_Destroy();
        write_pixels(argv[3], out.data(), n);
    } else if (cmd == "polar2x2" && argc == 5) {
        std::vector<uint32_t> in = read_pixels(argv[2]);
        if (in.size() != 644 * 364 || !Polar_Create()) return 1;
        std::vector<uint32_t> res(in.size());
        Polar_Blit_2x2(res.data(), in.data(), atoi(argv[3]) != 0);
        Polar_Destroy();
        write_pixels(argv[4], res.data(), res.size());
    } else if (cmd == "polar_blit_a" && argc == 6) {
        std::vector<uint32_t> d = read_pixels(argv[2]), s = read_pixels(argv[3]);
        if (d.size() != n || s.size() != n || !Polar_Create()) return 1;
        Polar_BlitA(d.data(), s.data(), atoi(argv[4]) != 0);
        Polar_Destroy();
        write_pixels(argv[5], d.data(), n);
    } else if (cmd == "tunnelscape" && argc == 5) {
        const int frames = atoi(argv[3]);
        if (frames < 1 || !Tunnelscape_Create()) return 1;
        std::vector<uint32_t> all((size_t)frames * n);
        for (int k = 0; k < frames; ++k)
            Tunnelscape_Draw(all.data() + (size_t)k * n, (float)atof(argv[2]), 1.f);
        Tunnelscape_Destroy();
        write_pixels(argv[4], all.data(), all.size());
    } else if (cmd == "testpattern" && argc == 3) {
        if (!FxBlitter_Create()) return 1;
        FxBlitter_DrawTestPattern(out.data());
        FxBlitter_Destroy();
        write_pixels(argv[2], out.data(), n);
    } else if (cmd == "lifecycle" && argc == 2) {
        // run with CKD_SHIM_ASSETS naming an empty directory: every Create
        // that loads an asset fails, and must give its reference back
        auto expect = [](int refs, const char *where) {
            if (ckd_shim_context_refs() != refs) {
                fprintf(stderr, "%s: %d context references, expected %d\n", where,
                        ckd_shim_context_refs(), refs);
                exit(1);
            }
        };
        expect(0, "start");
        if (Landscape_Create() || Tunnelscape_Create() || Shadertoy_Create() || Ball_Create() ||
            Twister_Create()) {
            fprintf(stderr, "a Create succeeded without its assets\n");
            return 1;
        }
        expect(0, "after failed Creates");
        g_failed = false; // the failures above reported through SetLastError
        if (!Polar_Create() || !Polar_Create()) {
            fprintf(stderr, "Polar_Create failed\n");
            return 1;
        }
        expect(1, "Polar_Create twice");
        Polar_Destroy();
        Polar_Destroy();
        expect(0, "Polar_Destroy twice");
        if (!FxBlitter_Create() || !BoxBlur_Create()) {
            fprintf(stderr, "FxBlitter_Create or BoxBlur_Create failed\n");
            return 1;
        }
        expect(2, "FxBlitter_Create + BoxBlur_Create");
        FxBlitter_DrawTestPattern(out.data());
        BoxBlur_Destroy();
        FxBlitter_Destroy();
        expect(0, "after their Destroys");
        std::vector<uint32_t> a(64, 0x01020304u), b(64, 0x10203040u);
        Add32(a.data(), b.data(), 64);
        expect(1, "a free function holds one reference");
        Add32(a.data(), b.data(), 64);
        expect(1, "and only one");
        if (a[0] != 0x21426384u) { // two saturating adds of b
            fprintf(stderr, "Add32: %08x\n", a[0]);
            return 1;
        }
        printf("ok\n");
    } else {
        return 2;
    }
    return g_failed ? 1 : 0;
}
