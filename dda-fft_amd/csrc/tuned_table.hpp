// Launch schedule per (lattice, right-hand sides): written by
// dda-fft_amd/tune.py from measurements on an MI355X.  x_rb: right-hand
// sides per X-pass tile; x_blocks: X-pass launch size up to which groups of
// right-hand sides are spread over blocks (1 << 30: each group its own
// block); graph: iterations replayed as captured HIP graphs; fuse: z- and
// y-forward transforms of an i-plane in one kernel.
#pragma once

namespace ddafft {

struct TunedEntry {
    int ni, nj, nl, nrhs, x_rb, x_blocks, graph, fuse;
};

// realizations in flight in ddafft_solve_batch for the batched Monte-Carlo
// workload (the call's width argument 0 stands for it)
constexpr int kTunedBatchWidth = 16;

constexpr TunedEntry kTuned[] = {
    {32, 32, 32, 2, 1, 2048, 0, 1},   // X pass 12.6 us, solve 0.881 ms
    {64, 64, 64, 2, 2, 1 << 30, 0, 0},   // X pass 46.8 us, solve 2.583 ms
    {128, 128, 128, 2, 2, 1 << 30, 0, 0},   // X pass 433.0 us, solve 8.857 ms
    {256, 256, 256, 2, 1, 1 << 30, 0, 0},   // X pass 5041.1 us, solve 91.992 ms
    {32, 32, 32, 16, 2, 8192, 0, 1},   // X pass 37.4 us, solve 2.343 ms
    {32, 32, 64, 16, 2, 1 << 30, 0, 1},   // X pass 72.2 us, solve 4.093 ms
    {64, 64, 64, 16, 2, 1 << 30, 0, 1},   // X pass 322.9 us, solve 15.907 ms
    {32, 32, 64, 2, 2, 1 << 30, 0, 1},   // X pass 18.5 us, solve 1.094 ms
    {64, 64, 64, 4, 2, 8192, 0, 0},   // X pass 95.1 us, solve 4.682 ms
    {128, 128, 128, 4, 2, 1 << 30, 0, 0},   // X pass 796.1 us, solve 17.004 ms
};

}  // namespace ddafft
