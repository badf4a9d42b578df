"""Algorithmic HBM bytes per launch of each step kernel.

Each figure counts the bytes a kernel must read and write at the internal
layouts, each array once: inputs it consumes, outputs it produces, the
Legendre tables once per launch and a shared basic state once per launch.
``s`` is the element size (8 for FP64, 4 for FP32); complex values are 2 s.
"""

MI355X_HBM_GBS = 8000.0          # vendor peak, GB/s
MI355X_FP64_TFLOPS = 78.6
MI355X_FP32_TFLOPS = 157.3
MI355X_BF16_TFLOPS = 2500.0      # dense matrix cores


def kernel_bytes(E, L=20, J=64, I=128, nmax=42, s=8, nbasic=1, nfrc=1, fused_tend_fft=False,
                 grid_fused=False):
    """``fused_tend_fft``: the tendency kernel with the forward FFT inside
    (its "tend" reads what k_tend reads and writes the Fourier columns k_fft
    writes; "fft" is then empty).  ``grid_fused``: the grid phase in one
    kernel (k_grid: "tend" reads the synthesised Fourier columns, the basic
    state, the forcing and ps', and writes ps' and the tendencies' Fourier
    columns; "ifft" and "fft" are empty)."""
    M = nmax + 1
    nspec = M * (M + 1) // 2
    Jh = J // 2
    EL = E * L
    c = 2 * s
    tables = 2 * nspec * Jh * s
    cin = 6 * EL + 6 * E           # Fourier/grid columns of the synthesised state
    cout = 9 * EL + 2 * E          # columns of the grid tendencies
    rb = 9 * L + 6                 # basic-state rows
    out = {
        "syn_level": (nspec * 4 * EL + M * J * 6 * EL) * c + tables,
        "syn_surf": (nspec * 3 * E + M * J * 6 * E) * c + tables,
        "ifft": M * J * cin * c + J * cin * I * s,
        "tend": (J * cin * I + nbasic * J * rb * I + nfrc * J * L * I
                 + 4 * E * J * I + J * cout * I) * s,
        "fft": J * cout * I * s + M * J * cout * c,
        "ana_level": M * J * 9 * EL * c + nspec * 4 * EL * c + tables,
        "ana_surf": M * J * 2 * E * c + nspec * 2 * E * c + tables,
        "step": (5 * nspec * 4 * EL + nspec * 2 * E + nspec * 3 * E) * c + nbasic * M * L * L * s,
    }
    if fused_tend_fft:
        out["tend"] += M * J * cout * c - J * cout * I * s
        out["fft"] = 0
    if grid_fused:
        out["tend"] = (M * J * cin * c + M * J * cout * c
                       + (nbasic * J * rb * I + nfrc * J * L * I + 4 * E * J * I) * s)
        out["ifft"] = out["fft"] = 0
    return out


def compulsory_bytes_per_member_step(L=20, J=64, I=128, s=8, per_member_basic=False, nmax=42):
    """The path's compulsory traffic per member-step: read + write two grid
    time levels and read the forcing, plus a per-member basic state."""
    b = 4 * (4 * L + 1) * J * I * s + L * J * I * s
    if per_member_basic:
        b += (9 * L * J * I + L + 4 * J * I + (nmax + 1) * L * L) * s
    return b
