"""Launch configurations of the fused 3D cycle kernels found by
``fsgh_autotune`` on an MI355X (Sedov blast 256^3, Morton order): per
kernel the workgroup size, the waves-per-SIMD floor and whether work
ranges are dealt per XCD.  ``SGHSolver.set_tuning(TABLE)`` applies them
without re-timing; ``SGHSolver.autotune()`` measures again on the caller's
own mesh.  Every configuration gives bit-identical results.

Measured (mean per launch over 50 cycles of the bench, default = 256
threads, no floor, XCD dealing in the autotuner's runs): div_extremes
0.324 ms with 512 threads in launch order (default 0.339 ms); vel_bc_pos
1.096 ms with 128 threads in launch order (default 1.279 ms: its LDS copy
of the block's corner forces limits blocks per CU); force 1.334 ms with 256
threads, a 3-wave floor (2 VGPRs spilled) and XCD dealing (default 1.463 ms;
1.347 ms with 128 threads); energy_state_div 1.344 ms with 128 threads, a
3-wave floor request and XCD dealing (1.349 ms with 256 threads in launch
order, default 1.401 ms).
"""

TABLE = {
    "div_extremes": {"block": 512, "min_waves": 1, "xcd_order": 0},
    "force": {"block": 256, "min_waves": 3, "xcd_order": 1},
    "vel_bc_pos": {"block": 128, "min_waves": 1, "xcd_order": 0},
    "energy_state_div": {"block": 128, "min_waves": 3, "xcd_order": 1},
}
