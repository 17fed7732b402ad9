/*
 * ref_common.h -- shared bits of the reference harness.
 *
 * TEST INFRASTRUCTURE ONLY.  Each h_*.cpp file #includes one unmodified
 * reference translation unit from the reference checkout and exposes a few
 * of its (often file-static) functions through extern "C" wrappers.  Nothing
 * is stubbed: the link uses --gc-sections, so functions that would need
 * SDL2 / DevIL / BASS / GNU Rocket are never linked in.
 */
#pragma once
#define REF_API extern "C" __attribute__((visibility("default")))
