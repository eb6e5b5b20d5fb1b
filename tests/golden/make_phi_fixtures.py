"""Extract the reference's committed NTT phi tables into small .npy fixtures.

Inputs (reference checkout, read as data):
  hw/memory_file/twiddle/NTT_CORE_ARCH_GF64/R2_PSI{8,16,32}/twd_phi_{fwd,bwd}_N2048_<psi>.mem
      ROMs written by ntt_gf64.sage gen_phi_rom (ntt_gf64.sage:685-724) for the cuts
      listed in gen_gf64_rom_twd.sh:9-20 (PSI8 -> [4,7], PSI16/PSI32 -> [5,6]).
  hw/module/number_theoretic_transform/module/ntt_core_gf64/module/ntt_core_gf64_phi/rtl/
      ntt_core_gf64_phi_phi_pkg.sv  written by gen_all_phi (ntt_gf64.sage:621-680).

Outputs (tests/golden/):
  rom_phi_{fwd,bwd}_N2048_cut{4_7,5_6}.npy   full 2048-entry phi list, re-assembled from the
      per-psi ROM files: rom_psi[2*i + r] = phi[i*COEF + 2*psi + r]
  pkg_phi_{fwd,bwd,bwdwdiv}_N{n}.npy  the 2n powers of w_2n (index = power), for n in PKG_N
Run once in the build container:  python3 tests/golden/make_phi_fixtures.py
"""
import os
import re
import sys

import numpy as np

REF = os.environ.get("LUTHE_REFERENCE", "/root/reference")
OUT = os.path.dirname(os.path.abspath(__file__))
ROM_DIR = os.path.join(REF, "hw/memory_file/twiddle/NTT_CORE_ARCH_GF64")
PKG = os.path.join(REF, "hw/module/number_theoretic_transform/module/ntt_core_gf64/module/"
                   "ntt_core_gf64_phi/rtl/ntt_core_gf64_phi_phi_pkg.sv")
PKG_N = (2048, 1024, 64, 32, 4)


def rom_table(psi_dir, direction, N, coef):
    phi = [None] * N
    for psi in range(coef // 2):
        path = os.path.join(ROM_DIR, psi_dir, f"twd_phi_{direction}_N{N}_{psi}.mem")
        vals = [int(l, 16) for l in open(path) if l.strip()]
        for pos, v in enumerate(vals):
            i, r = divmod(pos, 2)
            phi[i * coef + psi * 2 + r] = v
    assert all(v is not None for v in phi)
    return np.array(phi, dtype=np.uint64)


def pkg_tables():
    text = open(PKG).read()
    out = {}
    for m in re.finditer(r"localparam \[2\*(\d+)-1:0\]\[63:0\] NTT_GF64_(FWD|BWD|BWD_WDIV)_N\d+_PHI_L = \{(.*?)\};",
                         text, re.S):
        n, kind, body = int(m.group(1)), m.group(2), m.group(3)
        vals = [int(v, 16) for v in re.findall(r"64'h([0-9a-fA-F]+)", body)]
        assert len(vals) == 2 * n
        out[(kind, n)] = np.array(vals[::-1], dtype=np.uint64)  # text lists index 2n-1 first
    return out


def main():
    if not os.path.isdir(ROM_DIR):
        sys.exit(f"reference not found at {REF}")
    for psi_dir, coef, cut in (("R2_PSI8", 16, "4_7"), ("R2_PSI16", 32, "5_6")):
        for d in ("fwd", "bwd"):
            np.save(os.path.join(OUT, f"rom_phi_{d}_N2048_cut{cut}.npy"), rom_table(psi_dir, d, 2048, coef))
    # PSI32 holds the same [5,6] tables split over 32 files: check they agree.
    for d in ("fwd", "bwd"):
        assert np.array_equal(rom_table("R2_PSI32", d, 2048, 64), rom_table("R2_PSI16", d, 2048, 32))
    tabs = pkg_tables()
    for (kind, n), arr in tabs.items():
        if n in PKG_N:
            name = {"FWD": "fwd", "BWD": "bwd", "BWD_WDIV": "bwdwdiv"}[kind]
            np.save(os.path.join(OUT, f"pkg_phi_{name}_N{n}.npy"), arr)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
