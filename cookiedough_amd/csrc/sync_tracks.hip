// sync_tracks.hip -- GNU Rocket track playback on the host: the .track
// files of a demo's sync/ folder, their key interpolation, and the effect
// parameter structs filled from them.  Host code only; no GPU context is
// needed.
//
// File and name format (3rdparty/rocket-stripped/lib/device.c:26-75,
// 309-331): a track `name` lives in <dir>/sync/_<encoded name>.track, where
// every character other than [A-Za-z0-9._/] becomes '-' and two upper-case
// hex digits.  The file holds an int key count, then per key an int row, a
// float value and a one-byte interpolation type, packed, rows ascending.  A
// missing file is a track without keys (value 0).
//
// Interpolation (3rdparty/rocket-stripped/lib/track.c:29-60), in double:
// before the first key the first value, from the last key on the last
// value; in between, at key i with floor(row) in [row_i, row_{i+1}), the
// key's type: step (value_i), linear, smooth (t*t*(3-2t)) or ramp (t^2),
// with t = (row - row_i) / (row_{i+1} - row_i) and the value
// value_i + (value_{i+1} - value_i) * t, the difference taken in float.
#include "ckd_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#pragma clang fp contract(off)

struct ckd_sync {
    std::string dir;
    struct Key {
        int row;
        float value;
        int type;
    };
    struct Track {
        std::string name;
        std::vector<Key> keys;
    };
    std::vector<Track> tracks;
};

namespace {

bool path_char(char ch)
{
    return (ch >= 'a' && ch <= 'z') || (ch >= 'A' && ch <= 'Z') || (ch >= '0' && ch <= '9') ||
           ch == '.' || ch == '_' || ch == '/';
}

std::string track_path(const std::string &dir, const char *name)
{
    static const char kHex[] = "0123456789ABCDEF";
    std::string p = dir;
    if (!p.empty() && p.back() != '/') p += '/';
    p += "sync/_";
    for (const char *c = name; *c; ++c) {
        if (path_char(*c)) {
            p += *c;
        } else {
            const unsigned char u = (unsigned char)*c;
            p += '-';
            p += kHex[u >> 4];
            p += kHex[u & 15];
        }
    }
    return p + ".track";
}

// index of the last key with row <= irow, -1 before the first key
// (sync_find_key + key_idx_floor, track.c:62-80 and track.h:29-35)
int key_floor(const std::vector<ckd_sync::Key> &keys, int irow)
{
    int lo = 0, hi = (int)keys.size();
    while (lo < hi) {
        const int mi = (lo + hi) / 2;
        if (keys[mi].row < irow) lo = mi + 1;
        else if (keys[mi].row > irow) hi = mi;
        else return mi;
    }
    return lo - 1;
}

// each effect's tracks in the order of its ckd_*_params fields after `time`
// (the Rocket::AddTrack calls of code/shadertoy.cpp:113-171,
// code/ball.cpp:426-441, code/torus-twister.cpp:155-157)
const char *const kPlasma[] = {"plasma:Speed", "plasma:Hue", "plasma:Gamma", "plasma:Desaturation"};
const char *const kNautilus[] = {"nautilus:Roll", "nautilus:Hue", "nautilus:Speed", "nautilus:Desat",
                                 "nautilus:Blur"};
const char *const kSpikey[] = {
    "spike:Speed", "spike:Roll", "spike:Specular", "spike:Desaturation", "spike:Hue",
    "spike:Gamma", "distSpike:xOffs", "distSpike:yOffs", "distSpike:zOffs", "distSpike:Warmup",
    "closeSpike:xOffs", "closeSpike:yOffs", "closeSpike:zOffs", "closeSpike:zOffsScale",
    "closeSpike:NormalGrain", "closeSpike:Scale", "closeSpike:Rim", "closeSpike:AspectMul",
    "closeSpike:MixBlurMap", "closeSpike:MixBlur", "closeSpike:MixMapBlur",
    "closeSpike:MixBlurOpacity"};
const char *const kSinuses[] = {"sinusesTunnel:Specular", "sinusesTunnel:Roll", "sinusesTunnel:Speed",
                                "sinusesTunnel:OffsX", "sinusesTunnel:Gamma", "sinusesTunnel:Hue",
                                "sinusesTunnel:Desaturation"};
const char *const kLaura[] = {"laura:Speed", "laura:Yaw", "laura:Pitch", "laura:Roll", "laura:Hue",
                              "laura:Saturate"};
const char *const kFdTunnel[] = {
    "tunnel:Boxy", "tunnel:FlowerScale", "tunnel:FlowerFreq", "tunnel:FlowerPhase", "tunnel:Speed",
    "tunnel:Roll", "tunnel:Pitch", "tunnel:Radius", "tunnel:MulU", "tunnel:MulV",
    "tunnel:LitTiles", "tunnel:LitBlur", "tunnel:Fog1", "tunnel:Fog2"};
const char *const kBall[] = {
    "ball:Blur", "ball:Radius", "ball:RayLength", "ball:Spikes", "ball:HasBeams",
    "ball:BaseShapeIndex", "ball:Speed", "ball:BeamAttenuation", "ball:BeamAlphaMin",
    "ball:RotateOffsX", "ball:RotateOffsY", "ball:Beams1", "ball:Beams2", "ball:Beams3",
    "ball:BallLowBeams"};
const char *const kTwister[] = {"twister:Speed", "twister::ShearSpeed", "twister:Blur"};

struct EffectTracks {
    const char *effect;
    const char *const *names;
    unsigned count;
    size_t structSize;
};

#define CKD_FX(name, table, type) {name, table, sizeof(table) / sizeof(table[0]), sizeof(type)}
const EffectTracks kEffects[] = {
    CKD_FX("plasma", kPlasma, ckd_plasma_params),
    CKD_FX("nautilus", kNautilus, ckd_nautilus_params),
    CKD_FX("spikey", kSpikey, ckd_spikey_params),
    CKD_FX("sinuses", kSinuses, ckd_sinuses_params),
    CKD_FX("laura", kLaura, ckd_laura_params),
    CKD_FX("fdtunnel", kFdTunnel, ckd_fdtunnel_params),
    CKD_FX("ball", kBall, ckd_ball_params),
    CKD_FX("twister", kTwister, ckd_twister_params),
};
#undef CKD_FX

const EffectTracks *find_effect(const char *effect)
{
    for (const EffectTracks &e : kEffects)
        if (!strcmp(e.effect, effect)) return &e;
    return nullptr;
}

} // namespace

extern "C" {

ckd_status ckd_sync_open(const char *dir, ckd_sync **out)
{
    CKD_REQUIRE(dir && out, "ckd_sync_open: null argument");
    *out = new ckd_sync;
    (*out)->dir = dir;
    return CKD_OK;
}

void ckd_sync_close(ckd_sync *s) { delete s; }

ckd_status ckd_sync_track(ckd_sync *s, const char *name, int *index)
{
    CKD_REQUIRE(s && name && index, "ckd_sync_track: null argument");
    for (size_t i = 0; i < s->tracks.size(); ++i)
        if (s->tracks[i].name == name) {
            *index = (int)i;
            return CKD_OK;
        }
    ckd_sync::Track t;
    t.name = name;
    const std::string path = track_path(s->dir, name);
    if (FILE *f = fopen(path.c_str(), "rb")) {
        int32_t n = 0;
        bool ok = fread(&n, 4, 1, f) == 1 && n >= 0 && n <= (1 << 24);
        for (int32_t k = 0; ok && k < n; ++k) {
            unsigned char rec[9];
            ok = fread(rec, 1, 9, f) == 9;
            if (!ok) break;
            ckd_sync::Key key;
            memcpy(&key.row, rec, 4);
            memcpy(&key.value, rec + 4, 4);
            key.type = rec[8];
            ok = key.type <= 3 && (t.keys.empty() || t.keys.back().row < key.row);
            t.keys.push_back(key);
        }
        fclose(f);
        if (!ok)
            return ckd::fail(CKD_EINVAL, "ckd_sync_track: %s: not a track file (truncated, rows "
                                         "out of order or a key type above 3)", path.c_str());
    }
    s->tracks.push_back(std::move(t));
    *index = (int)s->tracks.size() - 1;
    return CKD_OK;
}

double ckd_sync_get(const ckd_sync *s, int index, double row)
{
    if (!s || index < 0 || index >= (int)s->tracks.size()) return 0.0;
    const std::vector<ckd_sync::Key> &k = s->tracks[index].keys;
    if (k.empty()) return 0.0;
    const int i = key_floor(k, (int)floor(row));
    if (i < 0) return k.front().value;
    if (i > (int)k.size() - 2) return k.back().value;
    if (k[i].type == 0) return k[i].value;
    double t = (row - k[i].row) / (k[i + 1].row - k[i].row);
    if (k[i].type == 2) t = t * t * (3 - 2 * t);
    else if (k[i].type == 3) t = pow(t, 2.0);
    // the value difference is a float subtraction, as in track.c
    const float dv = k[i + 1].value - k[i].value;
    return (double)k[i].value + (double)dv * t;
}

double ckd_sync_row(double seconds)
{
    // kRowRate of code/audio.cpp:18: 170 BPM played as 174, 16 rows a beat
    const double kRowRate = (170.0 / (60.0 * (170.0 / 174.0))) * 16.0;
    return seconds * kRowRate;
}

ckd_status ckd_effect_tracks(const char *effect, const char *const **names, unsigned *count)
{
    CKD_REQUIRE(effect && names && count, "ckd_effect_tracks: null argument");
    const EffectTracks *e = find_effect(effect);
    CKD_REQUIRE(e, "ckd_effect_tracks: unknown effect '%s'", effect);
    *names = e->names;
    *count = e->count;
    return CKD_OK;
}

ckd_status ckd_sync_params(ckd_sync *s, const char *effect, const double *rows, const float *times,
                           unsigned nFrames, void *params)
{
    CKD_REQUIRE(s && effect && (nFrames == 0 || (rows && times && params)),
                "ckd_sync_params: null argument");
    const EffectTracks *e = find_effect(effect);
    CKD_REQUIRE(e, "ckd_sync_params: unknown effect '%s'", effect);
    int idx[32];
    for (unsigned i = 0; i < e->count; ++i) {
        const ckd_status st = ckd_sync_track(s, e->names[i], &idx[i]);
        if (st != CKD_OK) return st;
    }
    float *out = static_cast<float *>(params);
    const unsigned stride = (unsigned)(e->structSize / sizeof(float));
    for (unsigned f = 0; f < nFrames; ++f) {
        float *p = out + (size_t)f * stride;
        p[0] = times[f];
        // Rocket::getf (code/rocket.h:23-25)
        for (unsigned i = 0; i < e->count; ++i) p[1 + i] = (float)ckd_sync_get(s, idx[i], rows[f]);
    }
    return CKD_OK;
}

} // extern "C"
