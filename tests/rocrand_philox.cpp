// Prints Philox4x32-10 outputs of rocRAND's philox4x32_10_engine (host side)
// for the counters the tests ask for: argv = seed stream blk [blk...].
// engine(seed, subsequence=stream, offset=4*blk).next4() is Philox with
// counter {lo(blk), hi(blk), lo(stream), hi(stream)} and key {lo(seed), hi(seed)}.
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>
#include <cstdlib>

int main(int argc, char **argv) {
    if (argc < 4) return 2;
    unsigned long long seed = strtoull(argv[1], nullptr, 0);
    unsigned long long stream = strtoull(argv[2], nullptr, 0);
    for (int i = 3; i < argc; ++i) {
        unsigned long long blk = strtoull(argv[i], nullptr, 0);
        rocrand_device::philox4x32_10_engine e(seed, stream, 4ull * blk);
        uint4 r = e.next4();
        printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
