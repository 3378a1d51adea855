/*
 * Counter-based SplitMix64 draws for the synthetic-input generator.
 *
 * Draw i of stream `seed` is the i-th output of a SplitMix64 generator whose
 * state starts at `seed` (state_i = seed + (i+1)*0x9E3779B97F4A7C15, then the
 * standard 30/27/31 xor-shift-multiply finaliser).  Because each draw is a
 * pure function of (seed, i), inputs are identical however many threads
 * generate them and in whatever order.  Kept in C because the bit operations
 * rely on wrapping unsigned 64-bit arithmetic.
 */
#include <math.h>
#include <stdint.h>

static inline uint64_t splitmix64_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Uniform in [0,1) with 53 random bits. */
double msod_rng_uniform(uint64_t seed, uint64_t i)
{
    return (double)(splitmix64_at(seed, i) >> 11) * 0x1.0p-53;
}

/* Standard normal by Box-Muller from draws 2i and 2i+1 of the stream. */
double msod_rng_normal(uint64_t seed, uint64_t i)
{
    const double two_pi = 6.283185307179586476925286766559;
    double u1 = 1.0 - msod_rng_uniform(seed, 2u * i);  /* (0,1] */
    double u2 = msod_rng_uniform(seed, 2u * i + 1u);
    return sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
}
