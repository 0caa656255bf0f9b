"""Launch configurations of the fused 3D cycle kernels found by
``fsgh_autotune`` on an MI355X (Sedov blast 256^3, Morton order): per
kernel the workgroup size, the waves-per-SIMD floor and how work ranges are
dealt to the 8 XCDs (``xcd_order``: 0 launch order, 1 one contiguous range
per XCD, k >= 2 chunks of k consecutive ranges dealt round-robin).
``SGHSolver.set_tuning(TABLE)`` applies them without re-timing;
``SGHSolver.autotune()`` measures again on the caller's own mesh.  Every
configuration gives bit-identical results.

Measured (mean per launch over 50 cycles of the bench, one box): force
1.300 ms with 256 threads, a 3-wave floor and chunks of 16 (1.337 ms with
one contiguous range per XCD; the autotuner's default variant 1.456 ms);
energy_state_div 1.269 ms with 128 threads, a 3-wave floor request and
chunks of 16 (1.332 ms contiguous per XCD); vel_bc_pos 1.073 ms with 128
threads and chunks of 4 (1.095 ms in launch order, default 1.278 ms: its
LDS copy of the block's corner forces limits blocks per CU); div_extremes
0.325 ms with 512 threads and chunks of 4 (default 0.338 ms).  Chunks of 8
or 32 for the element kernels measure the same within 0.2 %.
"""

TABLE = {
    "div_extremes": {"block": 512, "min_waves": 1, "xcd_order": 4},
    "force": {"block": 256, "min_waves": 3, "xcd_order": 16},
    "vel_bc_pos": {"block": 128, "min_waves": 1, "xcd_order": 4},
    "energy_state_div": {"block": 128, "min_waves": 3, "xcd_order": 16},
}
