/* A plain C99 consumer of include/kpcmi.h: what an R package's glue or any
 * other C caller compiles against.  It exercises the host-only entry points
 * (no GPU needed): default options, task seeds, the cache-control API,
 * argument validation with kpc_last_error, and the orientation defaults.
 * Prints one line per check and exits non-zero on the first failure. */
#include <stdio.h>
#include <string.h>

#include "kpcmi.h"

#define CHECK(cond, what)                                   \
  do {                                                      \
    if (!(cond)) {                                          \
      printf("FAIL %s (%s)\n", what, kpc_last_error());     \
      return 1;                                             \
    }                                                       \
    printf("ok %s\n", what);                                \
  } while (0)

int main(void) {
  kpc_options o;
  kpc_default_options(&o);
  CHECK(o.ci_method == KPC_DCC_GAMMA && o.alpha == 0.1 && o.num_col == 35 &&
            o.max_conditioning_size < 0 && o.perm_stream == KPC_PERM_COUNTER && o.use_cache == 0,
        "default options");

  const int S[2] = {2, 5};
  unsigned long long s1 = 0, s2 = 0;
  CHECK(kpc_task_seed(2024, KPC_DCC_PERM, 0, 7, S, 2, &s1) == 0 &&
            kpc_task_seed(2024, KPC_DCC_PERM, 7, 0, S, 2, &s2) == 0 && s1 == s2 && s1 != 0,
        "task seed is orientation free");

  kpc_cache_info ci;
  CHECK(kpc_cache_control(KPC_CACHE_RESET, -1, &ci) == 0 && ci.entries == 0 &&
            ci.capacity == 262144 && ci.target_capacity == 131072,
        "cache reset");
  CHECK(kpc_cache_control(KPC_CACHE_CONFIGURE, 0, &ci) < 0 &&
            strstr(kpc_last_error(), "1048576") != NULL,
        "cache capacity bound reported");
  const double data[12] = {1, 2, 3, 4, 5, 6, 2, 1, 4, 3, 6, 5};
  CHECK(kpc_cache_state(data, 6, 2, &ci) == 0 && strlen(ci.dataset_key) == 32 &&
            ci.dataset_entries == 0,
        "dataset cache state");

  kpc_result* r = NULL;
  CHECK(kpc_skeleton_run(NULL, 10, 3, &o, &r) < 0 && r == NULL &&
            strlen(kpc_last_error()) > 0,
        "null data rejected");
  o.alpha = 1.5;
  CHECK(kpc_skeleton_run(data, 6, 2, &o, &r) < 0 && strstr(kpc_last_error(), "alpha") != NULL,
        "alpha outside (0, 1) rejected");

  kpc_orient_options oo;
  kpc_default_orient_options(&oo);
  CHECK(oo.orient_collider == 1 && oo.solve_confl == 0 && oo.rules[0] && oo.rules[1] &&
            oo.rules[2] && oo.alpha == 0.1,
        "orientation defaults");
  const unsigned char asym[4] = {0, 1, 0, 0};
  const int slen[4] = {-1, -1, -1, -1};
  int sep[4 * 64];
  memset(sep, 0, sizeof(sep));
  kpc_orient_result* orr = NULL;
  CHECK(kpc_wanpdag_run(data, 6, 2, asym, slen, sep, &oo, &orr) < 0 && orr == NULL &&
            strstr(kpc_last_error(), "symmetric") != NULL,
        "asymmetric skeleton rejected");
  printf("all checks passed\n");
  return 0;
}
