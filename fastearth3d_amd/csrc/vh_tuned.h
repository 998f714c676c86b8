// vh_tuned.h -- launch and tile shapes of the hot kernels, as chosen by the
// project's autotuner (python -m fastearth3d_amd.autotune, which measures
// every candidate on the GPU and rewrites this file). A -D on the compiler
// line still overrides each value.
#pragma once
#ifndef VH_TSH_TB_SYNTH
#define VH_TSH_TB_SYNTH 3   // batches per block of the tensor synthesis
#endif
#ifndef VH_TSH_SU
#define VH_TSH_SU 4         // degrees per parity whose table values the synthesis loads together
#endif
#ifndef VH_DRIFT_RW
#define VH_DRIFT_RW 32      // right-hand sides (orders x re/im) per block of the drift solve
#endif
#ifndef VH_TCHUNK
#define VH_TCHUNK 108       // tensor batches per synthesis / row / analysis launch
#endif
#ifndef VH_TA_PF
#define VH_TA_PF 1          // k-steps of table fragments the tensor analysis keeps in flight
#endif
#ifndef VH_TA_JC
#define VH_TA_JC 48         // latitudes per LDS chunk of the tensor analysis
#endif
#ifndef VH_DRIFT_AK
#define VH_DRIFT_AK 8       // k-steps of the G_l fragment the drift's wavefront loads at a time
#endif
#ifndef VH_DRIFT_MINW
#define VH_DRIFT_MINW 1     // wavefronts per SIMD the drift's register budget must allow (4: two blocks per CU)
#endif
#ifndef VH_DRIFT_ONEPASS
#define VH_DRIFT_ONEPASS 1  // 1: the drift builds its forcing in one pass with LDS atomics
#endif
#ifndef VH_ROWS_RJ
#define VH_ROWS_RJ 8        // latitudes per block of the grouped row pass (one wavefront each)
#endif
#ifndef VH_ROWS_ADJ_RJ
#define VH_ROWS_ADJ_RJ 4    // latitudes per block of the grouped reverse row pass
#endif
#ifndef VH_ROWS_ADJ_MINW
#define VH_ROWS_ADJ_MINW 2  // wavefronts per SIMD the grouped reverse row pass's registers must allow
#endif
