// rocrand_kat.cpp -- prints rocRAND's philox4x32_10 output block for a given
// (key, counter), using the open rocRAND header installed with ROCm, so the
// oracle's own Philox (oracle.cpp) can be cross-checked against it.
// TEST INFRASTRUCTURE ONLY (host code; no GPU is touched).
//
// usage: rocrand_kat key_lo key_hi c0 c1 c2 c3   (hex)  -> 4 hex words
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
#include <cstdlib>

// ten_rounds(counter, key) is the engine's block function (protected).
struct KatEngine : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};

int main(int argc, char** argv) {
    if (argc != 7) return 2;
    unsigned int v[6];
    for (int i = 0; i < 6; ++i) v[i] = (unsigned int)std::strtoul(argv[i + 1], nullptr, 16);
    KatEngine eng;
    uint4 c;
    c.x = v[2]; c.y = v[3]; c.z = v[4]; c.w = v[5];
    uint2 k;
    k.x = v[0]; k.y = v[1];
    uint4 o = eng.block(c, k);
    std::printf("%08x %08x %08x %08x\n", o.x, o.y, o.z, o.w);
    return 0;
}
