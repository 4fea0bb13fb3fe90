"""GPU parity of the fused surrogate material update (cia_surrogate_update).

* The kernel against the float64 oracle (``oracle.mlp.material_update``, with the
  kernel's bf16 weights) over several calls in a row, each call from the same
  anchors: commit decisions, anchor strains, fallback flags and trip counts exactly,
  anchor ages to float32 rounding, stresses and latents within the bf16 rounding of
  the activations (2 % of the network's contribution).
* Inside the explicit dynamics loop: the stresses recorded at probe points equal the
  oracle driven by the recorded strain history of those points.
"""

import numpy as np
import pytest
import torch

from oracle import mlp as OM

pytestmark = pytest.mark.gpu


def _gate_dict(sm):
    g = sm.gate()
    out = {k: getattr(g, k) for k in OM.GATE_FIELDS}
    out["stiffness"] = np.asarray(list(g.stiffness))
    return out


def _anchor_to_oracle(st):
    return {"strain": st.strain.t().cpu().numpy().copy(),
            "inelastic": st.inelastic.t().cpu().numpy().copy(),
            "latent": st.latent.t().double().cpu().numpy().copy(),
            "age": st.age.double().cpu().numpy().copy(), "fallback": st.fallback.cpu().numpy() > 0,
            "current": st.current.t().cpu().numpy().copy()}


def _material(seed, **kw):
    from concrete_impact_ai_amd import surrogate as S

    layers = OM.init_params(seed)
    params = S.SurrogateParams(layers)
    # oracle weights: the bf16 values the kernel reads
    w = params.w_bf16.float().double().cpu().numpy()
    layers_bf = [(w[int(params.w_offsets[i]):int(params.w_offsets[i + 1])].reshape(params.w_shapes[i]), b)
                 for i, (_, b) in enumerate(params.layers_numpy())]
    a = np.random.default_rng(seed).standard_normal((6, 6))
    stiff = kw.pop("stiffness", 3e10) * (a @ a.T / 6.0 + np.eye(6))
    scales = S.SurrogateScales(1e-4, 100.0, 1.0, 5e7, d_eps_max=kw.pop("d_eps_max", 1.0),
                               rate_lo=-1.0, rate_hi=1.0, latent_max=kw.pop("latent_max", 1e9),
                               d_eps_commit=1.2e-4, stiffness=tuple(stiff.ravel()))
    return S.SurrogateMaterial("sur", params, scales, 2400.0, 4000.0, **kw), layers_bf


def _raw_update(sm, st, strain, dt, trips):
    """One cia_surrogate_update call through the C ABI (no fallback handling)."""
    import ctypes as C

    from concrete_impact_ai_amd import _lib
    from concrete_impact_ai_amd import surrogate as S
    from concrete_impact_ai_amd.materials import stream_handle

    def p(t):
        return C.c_void_p(t.data_ptr())

    stress = torch.empty_like(st.strain)
    nxt = S._SurrogateState(torch.empty_like(st.strain), torch.empty_like(st.inelastic),
                            torch.empty_like(st.latent), torch.empty_like(st.age),
                            torch.empty_like(st.fallback), 0, current=torch.empty_like(st.current))
    psi = torch.empty(strain.shape[1], dtype=torch.float64, device=strain.device)
    diss = torch.empty_like(psi)
    gate = sm.gate()
    _lib.check(_lib.lib().cia_surrogate_update(
        stream_handle(), strain.shape[1], p(strain), float(dt), p(sm.params.w_bf16), p(sm.params.b),
        C.byref(gate), p(st.strain), p(st.inelastic), p(st.latent), p(st.age), p(st.fallback), p(stress),
        p(nxt.strain), p(nxt.inelastic), p(nxt.latent), p(nxt.age), p(nxt.fallback), p(trips),
        p(st.current), p(nxt.current), p(psi), p(diss)),
        "cia_surrogate_update")
    return stress, nxt, psi, diss


@pytest.mark.parametrize("n", [1, 127, 1000, 4133])
def test_surrogate_update_kernel_matches_oracle(gpu, n):
    from concrete_impact_ai_amd import surrogate as S

    sm, layers = _material(11)
    rng = np.random.default_rng(n)
    st = S._SurrogateKernel(sm).initial_state(n)
    trips = torch.zeros(1, dtype=torch.int32, device=st.strain.device)
    st.inelastic.copy_(torch.as_tensor(rng.standard_normal((6, n)) * 1e-4, device=st.inelastic.device))
    st.latent.copy_(torch.as_tensor(rng.standard_normal((6, n)) * 0.5, device=st.latent.device))
    g = _gate_dict(sm)
    a = _anchor_to_oracle(st)
    eps = np.zeros((n, 6))
    total_trips = 0
    for call in range(4):
        d = rng.standard_normal((n, 6)) * 2.5e-5
        d[rng.random(n) < 0.05] *= 8.0  # a few increments beyond the envelope (1.25e-4)
        eps = eps + d
        dt = 10.0 ** rng.uniform(-7, -6)
        strain = torch.as_tensor(np.ascontiguousarray(eps.T), device=st.strain.device)
        r_stress, r_state, r_psi, r_diss = _raw_update(sm, st, strain, dt, trips)
        s_ref, a_ref, trip, n_new = OM.material_update(layers, g, eps, dt, a)
        total_trips += n_new
        nxt = _anchor_to_oracle(r_state)
        committed = a_ref["age"] == 0.0
        assert np.array_equal(nxt["age"] == 0.0, committed), call
        assert np.array_equal(nxt["strain"], a_ref["strain"]), call
        np.testing.assert_allclose(nxt["age"], a_ref["age"], rtol=2e-7)
        assert np.array_equal(nxt["fallback"], a_ref["fallback"]), call
        assert int(trips.item()) == total_trips
        # the network's inelastic increment of this call, and its bf16 tolerance
        C6 = g["stiffness"].reshape(6, 6)
        d_in = eps - a["inelastic"] - np.linalg.solve(C6, s_ref.T).T   # oracle ein' - anchor ein
        contrib = np.max(np.abs(d_in))
        got = r_stress.t().cpu().numpy()
        assert np.max(np.abs(got - s_ref)) <= 2e-2 * contrib * np.max(np.abs(C6)) * 6, call
        assert np.max(np.abs(nxt["inelastic"] - a_ref["inelastic"])) <= 2e-2 * contrib, call
        assert np.max(np.abs(nxt["latent"] - a_ref["latent"])) <= 2e-2 * max(
            1.0, np.max(np.abs(a_ref["latent"] - a["latent"]))), call
        # energies: the identities on the kernel's own outputs (round-off), and the
        # oracle's energies within the bf16 tolerance of the stress
        got_cur = r_state.current.t().cpu().numpy()
        prev_cur = st.current.t().cpu().numpy()
        psi_id, diss_id = OM.material_energies(eps, got, prev_cur, got_cur, dt)
        np.testing.assert_allclose(r_psi.cpu().numpy(), psi_id, rtol=1e-12, atol=1e-12 * np.max(np.abs(psi_id)))
        np.testing.assert_allclose(r_diss.cpu().numpy(), diss_id, rtol=1e-10,
                                   atol=1e-10 * np.max(np.abs(diss_id)))
        np.testing.assert_allclose(got_cur, eps - np.linalg.solve(C6, got.T).T, rtol=0,
                                   atol=1e-9 * np.max(np.abs(eps)))
        psi_ref, _ = OM.material_energies(eps, s_ref, a["current"], a_ref["current"], dt)
        assert np.max(np.abs(r_psi.cpu().numpy() - psi_ref)) <= 5e-2 * np.max(np.abs(psi_ref)), call
        # the oracle continues from the kernel's anchors (the same decisions, bf16 values)
        nxt["current"] = got_cur
        st, a = r_state, nxt
    assert total_trips > 0 or n < 100


def test_surrogate_material_in_explicit_loop_matches_oracle_at_probes(gpu):
    from concrete_impact_ai_amd import dynamics as D
    from concrete_impact_ai_amd import materials as M
    from concrete_impact_ai_amd.solvers import DirichletDofSet, TimeIntegrationSettings

    # an untrained network: a wide increment envelope keeps every point on the
    # network (no fallback) while the pulse moves the slab
    sm, layers = _material(12, probe_points=np.arange(0, 8 * 6 * 6 * 3, 37), d_eps_max=1e6,
                           stiffness=5e9)  # P-wave speed below the 4000 m/s of the CFL step
    nx, ny, nz = 6, 6, 3
    h = 0.01
    bundle = D.build_structured_box_bundle((nx * h, ny * h, nz * h), (nx, ny, nz), sm)
    load = D.build_half_sine_pressure_load(bundle, 1e5, 2e-5, patch=((2, 4), (2, 4)))
    bottom = bundle.mesh_info.boundary_groups["bottom"].nodes
    dofs = (3 * bottom[:, None] + np.arange(3)).ravel()
    n = bundle.mesh_info.dof_map.size
    dt = 0.5 * h / 4000.0
    ts = TimeIntegrationSettings("explicit", dt, 12, cfl_safety_factor=0.5)
    D.solve_material_explicit(bundle, ts, np.zeros(n), np.zeros(n),
                              DirichletDofSet(dofs, np.zeros(dofs.size)), load, "three_dimensional",
                              M.MaterialUpdateSettings(50, 1e-10, 1e-12, 1e-10), record_fields=False)
    strain, stress, dts = sm.probe_history()
    assert strain.shape[0] == 13 and np.max(np.abs(strain)) > 0.0
    g = _gate_dict(sm)
    a = OM.initial_anchor(strain.shape[1])
    ref = []
    for t in range(strain.shape[0]):
        s_ref, a, _, _ = OM.material_update(layers, g, strain[t], dts[t], a)
        ref.append(s_ref)
    ref = np.asarray(ref)
    assert np.max(np.abs(stress - ref)) <= 2e-2 * np.max(np.abs(ref))
    assert sm.stats["calls"] == 13 and sm.stats["gate_trips"] == 0
