"""Algorithmic bytes per cell of every kernel of the step (FP64 storage).

Definition (the measurement protocol of this project): the sum over a kernel's
distinct input and output fields of element size, per interior cell
N = itot*jtot*ktot.  Stencil re-reads, halos, the z ghost layers, constant tables
and scalars are not counted; masked or band-limited sweeps still count every cell;
a byte mask counts 1 byte.  Scratch that a kernel keeps for itself (the Thomas
multipliers) is not an input or output and is not counted.
"""

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, vendor figure
FP64_PEAK_TFLOPS = 78.6

# bytes per interior cell, FP64 fields (8 B) unless noted
KERNEL_BYTES_FP64 = {
    "k_tube": 8 + 8 + 1,               # R phi; W phi_c, mask
    "k_adv_stage1": 8 * 5 + 1,         # R phi0 (= phi_s), u, v, w, mask; W phi_1
    "k_adv_stage23": 8 * 6 + 1,        # R phi0, phi_s, u, v, w, mask; W phi_out
    "k_reinit_stage1": 8 * 2,          # R phi0 (= phi_s); W phi_n1
    "k_reinit_stage23": 8 * 3,         # R phi0, phi_s; W phi_out
    "k_materials": 8 * 3,              # R phi; W miu, rho
    "k_curvature": 8 * 2,              # R phi; W kappa
    "k_recurvature": 8 * 3,            # R phi, kappa; W kappa'
    "k_gfm": 8 * 9,                    # R phi, kappa', u*, v*, w*; W p, p_x, p_y, p_z
    "k_rfft_x_fwd": 8 * 2,             # R/W p
    "k_rfft_y_fwd": 8 * 2,
    "k_rfft_x_inv": 8 * 2,
    "k_rfft_y_inv": 8 * 2,
    "k_thomas": 8 * 2,                 # R/W p
    "k_momentum": 8 * 22,              # R u,v,w,miu,rho,duold x3,pnew,pold,p_xyz,p_xyz_old; W dudt x3, duold x3
    "k_correc": 8 * 10,                # R u* x3, p, p_x, p_y, p_z; W u, v, w
    "k_pressure_rotate": 8 * 4,        # R p, pnew; W pnew, pold
    "k_avg_vel": 8 * 6,                # R u, v, w; W ucc, vcc, wcc
}


def step_bytes_per_cell(calls):
    """Algorithmic bytes of one step given {kernel: launches per step}."""
    return sum(KERNEL_BYTES_FP64.get(k, 0) * n for k, n in calls.items())
