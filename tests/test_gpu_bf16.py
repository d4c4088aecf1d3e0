"""The BF16 form of the MLP (HDNNP_MIXED_BF16): the forward layers take
BF16 operands on v_mfma_f32_16x16x16_bf16 with FP32 accumulation, the
backward pass stays FP32 (neuralNetworks.f90:94-158).

- The kernel's atomic energies are those of BF16 arithmetic: the network of
  the oracle's weights evaluated in FP64 on the kernel's own (FP32) symmetry
  functions, with every forward operand (weights and layer inputs) rounded
  to BF16 to nearest even and the biases to FP32, agrees to FP32
  accumulation error (<= 2e-5 (1 + |E_i|)); the same network without the
  rounding differs by far more (so the test would see FP32 operands, and a
  wrong operand layout or a truncating rounding would fail it).
- dE/dG is the FP32 backward pass through the BF16 forward's
  pre-activations (FP32 weights), to 1e-5 of its largest entry.
- The force, energy and stress errors against FP64 are measured and
  printed; the bench reports them for the 49,152-atom cell and the
  4,096-structure batch. The form is opt-in: the FP32 MLP stays the mixed
  default because BF16's force error exceeds the 1e-5 Ha/bohr bound.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import PEROV_PARAMS

pytestmark = pytest.mark.gpu


def bf16(x):
    """FP32 values rounded to BF16, to nearest even (as the kernel's bf16_rne)"""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def _act(a, c):
    return np.tanh(c) if a == "t" else c


def _dact(a, c):
    return 1.0 - np.tanh(c) ** 2 if a == "t" else np.ones_like(c)


def _network(pot, z, x, round_forward):
    """E (without the atom energy) and dE/dx of one atom; forward operands
    BF16 (round_forward) or FP32, backward with FP32 weights"""
    cs = []
    y = x
    for l, (W, b) in enumerate(pot.layers[z]):       # W[nout, nin]
        Wq = bf16(W) if round_forward else f32(W)
        yq = bf16(y) if round_forward else f32(y)
        c = Wq @ yq + f32(b)
        cs.append(c)
        y = _act(pot.acts[l], c)
    g = np.ones(1)
    for l in range(len(pot.layers[z]) - 1, -1, -1):
        W, _ = pot.layers[z][l]
        g = f32(W).T @ (g * _dact(pot.acts[l], cs[l]))
    return float(y[0]), g


def test_bf16_forward_is_bf16_arithmetic(pkg, wl, oracle):
    s = wl.perovskite(2, 2, 2, seed=31)
    pot = oracle.read_potential(PEROV_PARAMS)
    h = pkg.parseRuNNerInput(PEROV_PARAMS, "mixed-bf16")
    h.energy_and_forces(s)
    G, D, Ei = h.symfunctions(s.nat)
    worst_bf, worst_fp = 0.0, 0.0
    for i in range(s.nat):
        z = int(s.Z[i])
        nsf = pot.layers[z][0][0].shape[1]
        x = G[i, :nsf]
        e_bf, g_bf = _network(pot, z, x, True)
        e_fp, _ = _network(pot, z, x, False)
        e_k = Ei[i] - pot.atom_energy.get(z, 0.0)
        worst_bf = max(worst_bf, abs(e_k - e_bf) / (1 + abs(e_bf)))
        worst_fp = max(worst_fp, abs(e_k - e_fp) / (1 + abs(e_fp)))
        np.testing.assert_allclose(D[i, :nsf], g_bf, rtol=0, atol=1e-5 * max(1e-3, np.abs(g_bf).max()))
    print(f"atomic energies vs BF16 arithmetic {worst_bf:.2e}, vs FP32 operands {worst_fp:.2e}", flush=True)
    assert worst_bf <= 2e-5
    assert worst_fp >= 20 * max(worst_bf, 1e-7)


@pytest.mark.parametrize("cells", [(4, 4, 4)])
def test_bf16_errors_against_fp64(pkg, wl, cells):
    s = wl.perovskite(*cells, seed=5)
    r = {p: pkg.parseRuNNerInput(PEROV_PARAMS, p).energy_and_forces(s) for p in ("fp64", "mixed", "mixed-bf16")}
    E0, F0, L0 = r["fp64"]
    out = {}
    for p in ("mixed", "mixed-bf16"):
        E, F, L = r[p]
        out[p] = {"force_rmse": float(np.sqrt(np.mean((F - F0) ** 2))), "dE_per_atom": abs(E - E0) / s.nat,
                  "stress": float(np.abs(pkg.dEdLat2Stress(s.lat, L) - pkg.dEdLat2Stress(s.lat, L0)).max())}
    print(f"perovskite {cells} ({s.nat} atoms) vs FP64: {out}", flush=True)
    assert out["mixed"]["force_rmse"] <= 1e-5
    assert np.isfinite(out["mixed-bf16"]["force_rmse"]) and out["mixed-bf16"]["force_rmse"] < 1e-2
