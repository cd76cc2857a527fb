// Kernel configuration per box size and precision (CfgDefault<N, T>, see CfgBase in
// box.hpp).  Written by `python -m dnsbox_amd.tune.autotune`; the notes record the
// measurements behind each entry.
#pragma once

namespace dnsbox {

// N=32 FP32 (state FP32): table entry step 0.0470 ms
//   z: 0.0242 ms/step (table 0.0242, 2 candidates)
//   y: 0.0267 ms/step (table 0.0268, 8 candidates)
//   x: 0.0160 ms/step (table 0.0191, 5 candidates)
//   -> CfgBase<32, 64, 8, 256, Plan<4, 2, 4>, Plan<2, 2, 8>, Plan<16>, 0, 0>
// N=48 FP32 (state FP32): table entry step 0.0780 ms
//   z: 0.0470 ms/step (table 0.0470, 1 candidates)
//   y: 0.0274 ms/step (table 0.0276, 5 candidates)
//   x: 0.0218 ms/step (table 0.0218, 7 candidates)
//   -> CfgBase<64, 64, 16, 256, Plan<2, 12, 2>, Plan<4, 12>, Plan<8, 3>, 0, 0>
// N=64 FP32 (state FP32): table entry step 0.0752 ms
//   z: 0.0321 ms/step (table 0.0378, 3 candidates)
//   y: 0.0303 ms/step (table 0.0303, 8 candidates)
//   x: 0.0264 ms/step (table 0.0272, 7 candidates)
//   -> CfgBase<16, 32, 16, 256, Plan<4, 4, 4>, Plan<8, 8>, Plan<16, 2>, 0, 0>
// N=128 FP64 (state FP64): table entry step 0.3353 ms
//   z: 0.1313 ms/step (table 0.1333, 2 candidates)
//   y: 0.1083 ms/step (table 0.1083, 8 candidates)
//   x: 0.1075 ms/step (table 0.1129, 4 candidates)
//   -> CfgBase<16, 16, 8, 256, Plan<8, 2, 8>, Plan<8, 4, 4>, Plan<16, 4>, 16, 6>
// N=256 FP32 (state FP64), passes pruned to the 2/3-rule cube: table entry step 1.0058 ms
//   z: 0.3217 ms/step (table 0.3217, 9 candidates)
//   y: 0.3224 ms/step (table 0.3224, 10 candidates)
//   x: 0.3393 ms/step (table 0.3736, 10 candidates)
//   -> CfgBase<16, 32, 8, 512, Plan<8, 4, 8>, Plan<16, 16>, Plan<16, 8>, 32, 2>
//   (before pruning: step 1.6263 ms, z 0.7347, y 0.4687, x 0.4127 with RB 4, 256 threads)
//   LDS rows measured beyond the bank model's ranking (x ms/step): XPD 64 / XPAD 1 0.3112-0.3135,
//   64 / 0 0.3112, 32 / 2 0.3399, 0 / 4 0.3443, 32 / 6 0.3500, 128 / 1 0.3425, 64 / 2 0.3674
//   -> CfgBase<16, 32, 8, 512, Plan<8, 4, 8>, Plan<16, 16>, Plan<16, 8>, 64, 1>
//   (256^3 FP64 and 512^3 FP32 keep their layouts: XPD 64 was 8 % and 6 % slower there)
// N=256 FP64 (state FP64), passes pruned to the 2/3-rule cube: table entry step 1.9779 ms
//   z: 0.6024 ms/step (table 0.6024, 3 candidates)
//   y: 0.6425 ms/step (table 0.6425, 10 candidates)
//   x: 0.5935 ms/step (table 0.7346, 7 candidates)
//   -> CfgBase<8, 16, 4, 256, Plan<8, 4, 8>, Plan<16, 16>, Plan<16, 8>, 32, 2>
//   (before pruning: step 2.8914 ms, z 1.1854, y 0.8418, x 0.7916 with RB 2, Plan<8, 4, 4>)
// N=512 FP32 (state FP64), passes pruned to the 2/3-rule cube: table entry step 7.9484 ms
//   z: 2.4014 ms/step (table 2.4014, 6 candidates)
//   y: 2.7309 ms/step (table 2.7309, 10 candidates)
//   x: 2.5708 ms/step (table 2.8147, 6 candidates)
//   -> CfgBase<8, 16, 4, 512, Plan<8, 8, 8>, Plan<16, 16, 2>, Plan<16, 16>, 32, 6>
//   (before pruning: step 13.6997 ms, z 6.6195, y 3.8948, x 3.1373 with RB 2, 256 threads)
// tuned 2026-12-03; N=256 FP32 / FP64 and N=512 FP32 re-tuned 2026-12-04 on the pruned step

template <> struct CfgDefault<16, float> : CfgBase<64, 64, 64, 256, Plan<4, 4>, Plan<4, 4>, Plan<8>, 0, 0> {};
template <> struct CfgDefault<32, float> : CfgBase<32, 64, 8, 256, Plan<4, 2, 4>, Plan<2, 2, 8>, Plan<16>, 0, 0> {};
template <> struct CfgDefault<48, float> : CfgBase<64, 64, 16, 256, Plan<2, 12, 2>, Plan<4, 12>, Plan<8, 3>, 0, 0> {};
template <> struct CfgDefault<64, float> : CfgBase<16, 32, 16, 256, Plan<4, 4, 4>, Plan<8, 8>, Plan<16, 2>, 0, 0> {};
template <> struct CfgDefault<128, float> : CfgBase<16, 32, 8, 256, Plan<4, 8, 4>, Plan<16, 8>, Plan<8, 8>, 16, 2> {};
template <> struct CfgDefault<256, float> : CfgBase<16, 32, 8, 512, Plan<8, 4, 8>, Plan<16, 16>, Plan<16, 8>, 64, 1> {};
template <> struct CfgDefault<512, float> : CfgBase<8, 16, 4, 512, Plan<8, 8, 8>, Plan<16, 16, 2>, Plan<16, 16>, 32, 6> {};
template <> struct CfgDefault<16, double> : CfgBase<64, 64, 32, 256, Plan<4, 4>, Plan<4, 4>, Plan<8>, 0, 0> {};
template <> struct CfgDefault<32, double> : CfgBase<32, 64, 16, 256, Plan<4, 2, 4>, Plan<8, 4>, Plan<4, 4>, 8, 0> {};
template <> struct CfgDefault<48, double> : CfgBase<64, 64, 8, 256, Plan<2, 12, 2>, Plan<12, 4>, Plan<8, 3>, 0, 10> {};
template <> struct CfgDefault<64, double> : CfgBase<16, 32, 8, 256, Plan<4, 4, 4>, Plan<8, 8>, Plan<8, 4>, 16, 0> {};
template <> struct CfgDefault<128, double> : CfgBase<16, 16, 8, 256, Plan<8, 2, 8>, Plan<8, 4, 4>, Plan<16, 4>, 16, 6> {};
template <> struct CfgDefault<256, double> : CfgBase<8, 16, 4, 256, Plan<8, 4, 8>, Plan<16, 16>, Plan<16, 8>, 32, 2> {};
template <> struct CfgDefault<512, double> : CfgBase<8, 8, 1, 256, Plan<8, 8, 8>, Plan<8, 8, 8>, Plan<8, 8, 4>, 16, 0> {};

}  // namespace dnsbox
