"""Write golden input/output vectors of the reference's native CPU ops.

The outputs come from the reference's own C++ (csrc/torch/*.cpp) compiled in
place into ``oracle/_ref`` by ``make -C oracle ref``; the vectors pin the C
restatement in ``oracle/grog_oracle.c``.  Run in the build container:

    python tests/golden/make_golden.py

Inputs are seeded; sorted indices (the reference's CPU scatter assumes them)
with one 4096+ case so the reference's threaded clean partition is used.
"""

from __future__ import annotations

import pathlib
import sys

import numpy as np
import torch

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent / "ref_ops.npz"


def main() -> None:
    ref = oracle.ref_ext()
    if ref is None:
        raise SystemExit("oracle/_ref is not built: run `make -C oracle ref`")
    torch.set_num_threads(4)
    rng = np.random.default_rng(20240601)
    out = {}
    for ct, rt, tag in ((np.complex64, np.float32, "c64"), (np.complex128, np.float64, "c128")):
        N, B, C, K = 96, 2, 4, 6
        d = (rng.standard_normal((N, B, C)) + 1j * rng.standard_normal((N, B, C))).astype(ct)
        k = (rng.standard_normal((K, C, C)) + 1j * rng.standard_normal((K, C, C))).astype(ct)
        idx = rng.integers(0, K, N).astype(np.int64)
        res = ref.grog_interpolate(torch.from_numpy(d), torch.from_numpy(idx), torch.from_numpy(k)).numpy()
        out.update({f"interp_{tag}_data": d, f"interp_{tag}_idx": idx, f"interp_{tag}_kernel": k,
                    f"interp_{tag}_out": res})
        for n, G in ((300, 40), (6000, 200)):
            si = np.sort(rng.integers(0, G, n)).astype(np.int64)
            w = rng.random(n).astype(rt)
            dat = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(ct)
            g0 = (rng.standard_normal(G) + 1j * rng.standard_normal(G)).astype(ct)
            g = torch.from_numpy(g0.copy())
            ref.scatter_add(g, torch.from_numpy(dat), torch.from_numpy(si), torch.from_numpy(w))
            gat = ref.gather(torch.from_numpy(g0), torch.from_numpy(si), torch.from_numpy(w)).numpy()
            p = torch.zeros(G, dtype=torch.float32 if rt == np.float32 else torch.float64)
            ref.psf_scatter_scalar(p, torch.from_numpy(si), torch.from_numpy(w))
            key = f"{tag}_{n}"
            out.update({f"sc_{key}_grid0": g0, f"sc_{key}_data": dat, f"sc_{key}_idx": si, f"sc_{key}_w": w,
                        f"sc_{key}_out": g.numpy(), f"ga_{key}_out": gat, f"psf_{key}_out": p.numpy()})
        M, n, G = 3, 500, 30
        si = np.sort(rng.integers(0, G, n)).astype(np.int64)
        w = rng.random(n).astype(rt)
        b = (rng.standard_normal((n, M)) + 1j * rng.standard_normal((n, M))).astype(ct)
        p = torch.zeros(G, M, M, dtype=torch.from_numpy(b).dtype)
        ref.psf_scatter_outer(p, torch.from_numpy(si), torch.from_numpy(w), torch.from_numpy(b))
        out.update({f"po_{tag}_idx": si, f"po_{tag}_w": w, f"po_{tag}_b": b, f"po_{tag}_out": p.numpy()})
        Kb, T = 3, 7
        bs = (rng.standard_normal((Kb, T)) + 1j * rng.standard_normal((Kb, T))).astype(ct)
        ti = rng.integers(-1, T + 1, n).astype(np.int64)  # includes out-of-range times
        p = torch.zeros(G, Kb, Kb, dtype=torch.from_numpy(bs).dtype)
        ref.psf_scatter_outer_basis(p, torch.from_numpy(si), torch.from_numpy(w), torch.from_numpy(bs),
                                    torch.from_numpy(ti))
        out.update({f"pb_{tag}_basis": bs, f"pb_{tag}_t": ti, f"pb_{tag}_out": p.numpy()})
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
