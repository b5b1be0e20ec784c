"""Algorithmic bytes per cell of every kernel of the step.

Definition (the measurement protocol of this project): the sum over a kernel's
distinct input and output arrays of their element size, per interior cell
N = itot*jtot*ktot.  Stencil re-reads, halos, the z ghost layers, constant tables
and scalars are not counted; masked or band-limited sweeps still count every cell;
a byte mask counts 1 byte.  Scratch that a kernel keeps for itself (the Thomas
multipliers and forward-sweep values of the serial Thomas kernel, used where the
partitioned one does not fit) is not an input or output and is not counted.

Arrays are of three classes: "T" field arrays (8 bytes in the FP64 mode, 4 in the
mixed modes), "P" pressure arrays (8 bytes in every mode) and "B" byte masks.
The face-jump fields p_x .. p_zold are read and written only where their byte mask
says they may be nonzero (a few cells around the interface); by the rule above they
still count every cell, so GFM, momentum and correc move fewer bytes than counted.
"""

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, vendor figure
FP64_PEAK_TFLOPS = 78.6

# kernel -> (number of T arrays, number of P arrays, number of byte masks) read or written
KERNEL_ARRAYS = {
    "k_tube": (2, 0, 1),               # R phi; W phi_c, mask
    "k_adv_stage1": (5, 0, 1),         # R phi0 (= phi_s), u, v, w, mask; W phi_1
    "k_adv_stage23": (6, 0, 1),        # R phi0, phi_s, u, v, w, mask; W phi_out
    "k_reinit_stage1": (2, 0, 0),      # R phi0 (= phi_s); W phi_n1
    "k_reinit_stage23": (3, 0, 0),     # R phi0, phi_s; W phi_out
    "k_materials": (3, 0, 0),          # R phi; W miu, rho
    "k_curvature": (2, 0, 0),          # R phi; W kappa
    "k_recurvature": (3, 0, 0),        # R phi, kappa; W kappa'
    "k_gfm": (9, 0, 1),                # R phi, kappa', u*, v*, w*; W rhs, p_x, p_y, p_z, jump mask
    "k_momentum": (20, 2, 2),          # R u,v,w,miu,rho,duold x3,p_xyz,p_xyz_old; R pnew,pold; W dudt x3, duold x3; R 2 jump masks
    "k_correc": (9, 1, 1),             # R u* x3, p_x, p_y, p_z; R p; W u, v, w; R jump mask
    "k_pressure_rotate": (0, 4, 0),    # R p, pnew; W pnew, pold
    "k_avg_vel": (6, 0, 0),            # R u, v, w; W ucc, vcc, wcc
    "k_residual": (2, 1, 0),           # R rhs, p; W r (mixed modes only)
}

# the Poisson passes differ by mode: FP64 works in place on p (P class); the mixed
# passes read and write the float work array.  Inside a mixed time step the one
# refinement pass starts from the previous step's solution rounded to float, p0: its
# forward x transform forms rhs - L p0 itself (R rhs, p0; W w: 12) and its inverse x
# transform writes p0 + e as double and, rounded, back into p0 (R w, p0; W p, p0: 20).
# Where no previous solution is held (the first step, an API solve) a float pass comes
# first: its forward x transform reads rhs and writes w (8) and its inverse writes p0
# (8); those launches are not in the steady-state step.  k_residual runs only for line
# lengths without a fixed-size transform.
_POISSON_FP64 = {"k_rfft_x_fwd": 16, "k_rfft_y_fwd": 16, "k_rfft_y_inv": 16, "k_rfft_x_inv": 16, "k_thomas": 16}
_POISSON_MIXED = {"k_rfft_x_fwd": 8, "k_rfft_x_fwd_res": 12, "k_rfft_y_fwd": 8, "k_rfft_y_inv": 8,
                  "k_rfft_x_inv": 20, "k_thomas": 8}


def kernel_bytes(mode_fp32):
    """{kernel: algorithmic bytes per interior cell} for the FP64 (False) or a mixed
    (True) mode."""
    t = 4 if mode_fp32 else 8
    out = {k: nt * t + npr * 8 + nb for k, (nt, npr, nb) in KERNEL_ARRAYS.items()}
    out.update(_POISSON_MIXED if mode_fp32 else _POISSON_FP64)
    if not mode_fp32:
        del out["k_residual"]
    return out


KERNEL_BYTES_FP64 = kernel_bytes(False)
KERNEL_BYTES_MIXED = kernel_bytes(True)


def step_bytes_per_cell(calls, mode_fp32=False):
    """Algorithmic bytes of one step given {kernel: launches per step}."""
    kb = kernel_bytes(mode_fp32)
    return sum(kb.get(k, 0) * n for k, n in calls.items())


# profiler scope name -> the kernel symbol (template arguments as the profiler prints
# them; {T} is float or double) whose HBM traffic profiles/traffic.json records
_TRAFFIC_KERNEL = {
    "k_momentum": "k_momentum_zm<{T}, 0, true, true>",   # power-of-two grid, 32-bit offsets
    "k_correc": "k_correc<{T}>",
    "k_gfm": ("k_gfm4<float>", "k_gfm<double, 2>"),     # four cells per thread in float only
    "k_tube": "k_tube_zm<{T}, 2>",
    "k_reinit_stage1": "k_reinit_zm<{T}, 1, true, 4, true>",
    "k_reinit_stage23": "k_reinit_zm<{T}, 3, true, 4, true>",
    "k_materials": "k_materials1<{T}>",                  # one level set on the whole grid
    "k_curvature": "k_curvature4<{T}>",
    "k_thomas": "k_thomas_seg<{T}, 16>",
}


def traffic_key(scope, fp32):
    """The traffic.json key of the kernel a profiler scope times, or None."""
    k = _TRAFFIC_KERNEL.get(scope)
    if k is None:
        return None
    if isinstance(k, tuple):
        return k[0] if fp32 else k[1]
    return k.format(T="float" if fp32 else "double")

