/* Harness over the reference's own vendored wyhash (include/wyhash/wyhash.h,
 * "final version 4.2"), compiled from the file where it lies in the reference
 * tree.  Built into oracle/_ref/ only; used to write golden vectors that pin the
 * oracle's restatement of wyhash.  The salt is Cuttlefish's wyhash salt
 * (include/Minimizer_Utility.hpp:26, include/Kmer.hpp:43). */
#include <stddef.h>
#include <stdint.h>
#include "wyhash/wyhash.h"

static const uint64_t cf_salt[4] = {4167021922371662411ull, 7320285940802167691ull,
                                    14307255741305819987ull, 10859488101230029397ull};

uint64_t ref_wyhash(const void* key, size_t len, uint64_t seed) { return wyhash(key, len, seed, cf_salt); }

/* Minimizer_Utility::hash in a release build: wyhash(&lmer, 8, seed, salt). */
uint64_t ref_lmer_hash(uint64_t lmer, uint64_t seed) { return wyhash(&lmer, sizeof(lmer), seed, cf_salt); }
