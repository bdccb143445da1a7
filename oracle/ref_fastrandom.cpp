// ref_fastrandom.cpp -- exposes the reference's own FastRandom
// (libs/utils/libutils/fast_random.h:8-40, compiled from where it lies in the
// clBool tree) through a C entry point, so the oracle's restatement can be
// checked draw for draw.  TEST INFRASTRUCTURE ONLY: built into oracle/_ref by
// oracle/Makefile and loaded only by tests.
#include <cstdint>
#include <libutils/fast_random.h>

extern "C" void ref_fastrandom_stream(uint64_t seed, int32_t min, int32_t max, uint32_t n, int32_t *out) {
    FastRandom r(static_cast<unsigned long>(seed));
    for (uint32_t i = 0; i < n; ++i) out[i] = r.next(min, max);
}

extern "C" int32_t ref_unsigned_long_bits(void) { return static_cast<int32_t>(sizeof(unsigned long) * 8); }
