/*
 * sharp_announce.h -- start-up banner of libsharp's programs.
 * Same declarations as libsharp/sharp_announce.h:35-37; the banner names the
 * GPU engine instead of libsharp's CPU vector length and OpenMP threads.
 */
#ifndef PLANCK_SHARP_ANNOUNCE_H
#define PLANCK_SHARP_ANNOUNCE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Prints `name` in a frame, then the engine and the visible GPUs
   (libsharp/sharp_announce.c:75-89). */
void sharp_announce (const char *name);

/* Banner when verbose; if argc != argc_expected, prints
   "Usage: <name> <argv_expected>" to stderr and exits with status 1
   (libsharp/sharp_announce.c:91-98). */
void sharp_module_startup (const char *name, int argc, int argc_expected,
  const char *argv_expected, int verbose);

#ifdef __cplusplus
}
#endif

#endif
