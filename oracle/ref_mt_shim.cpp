// C-linkage entry points over the reference's own MT19937
// (src/mt19937ar.cpp:57 init_genrand, :102 genrand_int32), compiled from the
// reference sources where they lie.  Test infrastructure only.
#include <stdint.h>
void init_genrand(unsigned long s);
unsigned long genrand_int32(void);
extern "C" void ref_mt_fill(uint32_t seed, int64_t n, uint32_t *out) {
    init_genrand(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = (uint32_t)genrand_int32();
}
