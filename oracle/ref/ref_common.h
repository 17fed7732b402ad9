/*
 * ref_common.h -- shared bits of the reference harness.
 *
 * TEST INFRASTRUCTURE ONLY.  Each h_*.cpp file #includes one unmodified
 * reference translation unit from the reference checkout and exposes a few
 * of its (often file-static) functions through extern "C" wrappers.  Nothing
 * is stubbed: the link uses --gc-sections, so functions that would need
 * SDL2 / DevIL / BASS are never linked in.  GNU Rocket is the reference's
 * own player library (3rdparty/rocket-stripped/lib, C sources), built here
 * in SYNC_PLAYER mode so the effects read their tracks from .track files.
 */
#pragma once
#define REF_API extern "C" __attribute__((visibility("default")))
