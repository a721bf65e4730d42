"""Writes the oracle's golden fixtures: tests/golden/oracle_golden.json (case
list, SHA-256 of the FP64 bytes of each output) and oracle_golden.npz (the
full arrays of two cases).  Inputs are closed-form (a Taylor-Green vortex with
a deterministic trigonometric perturbation), so no random generator is
involved.  Outputs: the right-hand side of the input state (interior cells)
and the state QJ after 3 RK steps (ghost layers included).

Run from the repository root with the oracle built (make -C oracle):
    python tests/golden/make_golden.py
The fixtures pin the oracle against unintended change; the GPU library is
checked against the same digests in tests/test_gpu_golden.py.
"""
import hashlib
import json
import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

from ouxsbli_amd import cases  # noqa: E402

# (order, scheme, visc, visc_order, tvd, slau_variant, rk[, viscous stencil])
CASES = [
    (2, "KEEP", "Euler", 0, "none", "HRSLAU2", 4),
    (4, "KEEP", "NS", 4, "none", "HRSLAU2", 4),
    (6, "KEEP", "NS", 6, "none", "HRSLAU2", 3),
    (4, "SLAU", "Euler", 0, "tvd", "HRSLAU2", 3),
    (6, "SLAU", "NS", 6, "hybrid", "HRSLAU2", 3),
    (6, "SLAU", "NS", 4, "hybrid", "SLAU", 4),
    (2, "Roe", "Euler", 0, "none", "HRSLAU2", 3),
    (6, "Roe", "NS", 6, "hybrid", "HRSLAU2", 3),
    (4, "Hybrid", "NS", 4, "tvd", "HRSLAU2", 4),
    (6, "Hybrid", "LES", 6, "hybrid", "HRSLAU2", 4),
    (6, "SLAU", "LES", 6, "hybrid", "HRSLAU2", 3),
    (6, "SLAU", "NS", 6, "hybrid", "HRSLAU2", 3, "ME4Base"),
    (6, "Hybrid", "LES", 6, "hybrid", "HRSLAU2", 4, "ME4Base"),
]
FULL = {(4, "KEEP", "NS", 4, "none", "HRSLAU2", 4), (6, "SLAU", "NS", 6, "hybrid", "HRSLAU2", 3)}
N = (12, 14, 13)  # each axis >= 2*ORDER, as the library requires
STEPS = 3


def case_id(c):
    return "-".join(map(str, c))


def spec_of(c):
    order, scheme, visc, vo, tvd, slau, rk = c[:7]
    stencil = c[7] if len(c) > 7 else "Central"
    s = cases.tgv_spec(N[0], order, Ma=0.8, cfl=0.3, cfl_mode="acoustic", scheme=scheme, visc=visc,
                       visc_order=vo, visc_stencil=stencil, tvd=tvd, slau_variant=slau, rk=rk)
    s.ny, s.nz = N[1] + order, N[2] + order
    s.Ly, s.Lz = s.Lx * N[1] / N[0], s.Lx * N[2] / N[0]
    return s


def input_state(spec):
    """TGV plus a deterministic perturbation that breaks its symmetries."""
    Q = cases.tgv_state(spec)
    x = cases.cell_centres(spec.nx, spec.Lx, spec.order)
    y = cases.cell_centres(spec.ny, spec.Ly, spec.order)
    z = cases.cell_centres(spec.nz, spec.Lz, spec.order)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    f = np.sin(3 * X + 1.0) * np.cos(2 * Y + 0.5) * np.sin(Z + 0.3)
    g = np.cos(X + 0.2) * np.sin(3 * Y + 0.7) * np.cos(2 * Z + 0.1)
    for l, (a, h) in enumerate([(0.02, f), (0.05, g), (0.05, f), (0.05, g), (0.01, f + g)]):
        scale = np.abs(Q[:, :, l, :]).max() if l != 3 else np.abs(Q[:, :, 1, :]).max()
        Q[:, :, l, :] = Q[:, :, l, :] + a * scale * h
    return cases.periodic_fill(Q, spec.order)


def interior(spec, A):
    h = spec.order // 2
    return np.ascontiguousarray(A[h:spec.nz - h, h:spec.ny - h, :, h:spec.nx - h])


def digest(A):
    return hashlib.sha256(np.ascontiguousarray(A, dtype="<f8").tobytes()).hexdigest()


def main():
    from oracle_lib import Oracle
    out, arrays = [], {}
    for c in CASES:
        spec = spec_of(c)
        Q = input_state(spec)
        o = Oracle(spec)
        o.set_state(Q)
        R = interior(spec, o.rhs())
        o.step(STEPS)
        S = o.get_qj()
        out.append({"case": case_id(c), "rhs_sha256": digest(R), "state_sha256": digest(S),
                    "rhs_absmax": float(np.abs(R).max()), "state_absmax": float(np.abs(S).max())})
        if c in FULL:
            arrays[case_id(c) + "/input"] = Q
            arrays[case_id(c) + "/rhs"] = R
            arrays[case_id(c) + "/state"] = S
    meta = {"interior": N, "steps": STEPS, "cases": out}
    (HERE / "oracle_golden.json").write_text(json.dumps(meta, indent=1) + "\n")
    np.savez_compressed(HERE / "oracle_golden.npz", **arrays)
    print(f"wrote {len(out)} cases, {len(arrays)} arrays")


if __name__ == "__main__":
    main()
