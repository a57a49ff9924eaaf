/* A C caller of libtgrav.so through include/tgrav.h alone, the way a C or Fortran host code
 * (the FLASH side of the bridge) would bind it.
 *
 *   abi_drive MODE      (MODE 0 = TGRAV_PLAIN_FP64, 1 = TGRAV_MIXED)
 *
 * Checks the error paths that need no device, then evaluates a seeded problem with host
 * pointers: the field at explicit points (tgrav_eval), the same with the fused kick
 * (tgrav_eval_kick) and at leaf-block targets (tgrav_eval_blocks).  Prints every input and
 * output with %.17g (exact round trip) for the test that compares them with the oracle:
 *   S x y z m eps          one line per source
 *   T x y z ax ay az phi   one line per explicit target
 *   K vx vy vz             the kicked velocities (from 0, dt = 0.25)
 *   B ox oy oz dx          one line per leaf block, then
 *   C ax ay az phi         one line per cell, in the blocks' C order
 * Exit status 0 on success; a failing call prints tgrav_last_error() and exits non-zero. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "tgrav.h"

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;

static double urand(void) { /* xorshift64, uniform in [0, 1) */
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}

#define CHECK(call)                                                          \
  do {                                                                       \
    int rc_ = (call);                                                        \
    if (rc_ != TGRAV_OK) {                                                   \
      fprintf(stderr, "%s -> %d: %s\n", #call, rc_, tgrav_last_error());     \
      return 1;                                                              \
    }                                                                        \
  } while (0)

#define EXPECT_ERR(call)                                                     \
  do {                                                                       \
    if ((call) == TGRAV_OK) {                                                \
      fprintf(stderr, "%s succeeded, expected an error\n", #call);           \
      return 1;                                                              \
    }                                                                        \
  } while (0)

int main(int argc, char** argv) {
  const int mode = argc > 1 ? atoi(argv[1]) : TGRAV_PLAIN_FP64;
  enum { NS = 777, NT = 1500, NB = 3, NC = NB * 512 };
  static double sx[NS], sy[NS], sz[NS], sm[NS], se[NS];
  static double tx[NT], ty[NT], tz[NT], ax[NT], ay[NT], az[NT], phi[NT];
  static double vx[NT], vy[NT], vz[NT], kx[NT], ky[NT], kz[NT];
  static double ox[NB], oy[NB], oz[NB], bdx[NB];
  static double cx[NC], cy[NC], cz[NC], cphi[NC];
  tgrav_ctx* ctx = NULL;
  int64_t i;

  for (i = 0; i < NS; ++i) {
    sx[i] = 2.0 * urand() - 1.0;
    sy[i] = 2.0 * urand() - 1.0;
    sz[i] = 2.0 * urand() - 1.0;
    sm[i] = 0.1 + 1.9 * urand();
    se[i] = 0.001 + 0.049 * urand();
  }
  for (i = 0; i < NT; ++i) {
    tx[i] = 3.0 * urand() - 1.5;
    ty[i] = 3.0 * urand() - 1.5;
    tz[i] = 3.0 * urand() - 1.5;
    vx[i] = vy[i] = vz[i] = 0.0;
  }
  for (i = 0; i < NB; ++i) {
    ox[i] = urand() - 0.5;
    oy[i] = urand() - 0.5;
    oz[i] = urand() - 0.5;
    bdx[i] = 0.01 + 0.05 * urand();
  }

  /* error paths: no device work needed */
  EXPECT_ERR(tgrav_create(NULL, 0, NULL));
  EXPECT_ERR(tgrav_eval(NULL, NT, NULL, tx, ty, tz, ax, ay, az, phi, mode, TGRAV_HOST));
  EXPECT_ERR(tgrav_set_sources(NULL, NS, sx, sy, sz, sm, se, TGRAV_HOST));

  CHECK(tgrav_create(&ctx, 0, NULL));
  EXPECT_ERR(tgrav_eval(ctx, NT, NULL, tx, ty, tz, ax, ay, az, phi, mode, TGRAV_HOST)); /* no sources */
  EXPECT_ERR(tgrav_set_sources(ctx, -1, sx, sy, sz, sm, se, TGRAV_HOST));
  EXPECT_ERR(tgrav_eval(ctx, NT, NULL, tx, ty, tz, ax, ay, az, phi, 7, TGRAV_HOST));       /* bad mode */
  CHECK(tgrav_set_sources(ctx, NS, sx, sy, sz, sm, se, TGRAV_HOST));
  CHECK(tgrav_eval(ctx, NT, NULL, tx, ty, tz, ax, ay, az, phi, mode, TGRAV_HOST));
  CHECK(tgrav_eval_kick(ctx, NT, NULL, tx, ty, tz, 0.25, vx, vy, vz, kx, ky, kz, NULL, mode,
                        TGRAV_HOST));
  CHECK(tgrav_eval_blocks(ctx, NB, 8, ox, oy, oz, bdx, NULL, cx, cy, cz, cphi, 0.0, NULL, NULL,
                          NULL, mode, TGRAV_HOST));
  CHECK(tgrav_destroy(ctx));

  for (i = 0; i < NT; ++i)
    if (kx[i] != ax[i] || ky[i] != ay[i] || kz[i] != az[i]) {
      fprintf(stderr, "eval_kick acceleration differs from eval at %lld\n", (long long)i);
      return 1;
    }
  for (i = 0; i < NS; ++i)
    printf("S %.17g %.17g %.17g %.17g %.17g\n", sx[i], sy[i], sz[i], sm[i], se[i]);
  for (i = 0; i < NT; ++i)
    printf("T %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", tx[i], ty[i], tz[i], ax[i], ay[i],
           az[i], phi[i]);
  for (i = 0; i < NT; ++i) printf("K %.17g %.17g %.17g\n", vx[i], vy[i], vz[i]);
  for (i = 0; i < NB; ++i) printf("B %.17g %.17g %.17g %.17g\n", ox[i], oy[i], oz[i], bdx[i]);
  for (i = 0; i < NC; ++i) printf("C %.17g %.17g %.17g %.17g\n", cx[i], cy[i], cz[i], cphi[i]);
  return 0;
}
