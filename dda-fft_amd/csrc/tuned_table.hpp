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
constexpr int kTunedBatchWidth = 12;

constexpr TunedEntry kTuned[] = {
    {32, 32, 32, 2, 1, 2048, 0, 0},   // X pass 11.8 us, solve 0.893 ms
    {64, 64, 64, 2, 2, 1 << 30, 0, 0},   // X pass 42.4 us, solve 2.533 ms
    {128, 128, 128, 2, 2, 1 << 30, 1, 0},   // X pass 380.2 us, solve 8.411 ms
    {256, 256, 256, 2, 2, 1 << 30, 1, 0},   // X pass 4001.6 us, solve 83.154 ms
    {32, 32, 32, 16, 2, 1 << 30, 0, 1},   // X pass 34.5 us, solve 2.210 ms
    {32, 32, 64, 16, 2, 1 << 30, 0, 1},   // X pass 69.2 us, solve 4.040 ms
    {64, 64, 64, 16, 2, 1 << 30, 1, 1},   // X pass 295.1 us, solve 15.458 ms
    {32, 32, 64, 2, 2, 1 << 30, 0, 1},   // X pass 16.7 us, solve 1.055 ms
    {64, 64, 64, 4, 2, 1 << 30, 0, 0},   // X pass 90.6 us, solve 4.521 ms
    {128, 128, 128, 4, 2, 1 << 30, 0, 0},   // X pass 708.0 us, solve 16.209 ms
};

}  // namespace ddafft
