"""The fused field + LLG RHS kernel's fast form against its general form.

smx_field_llg_rhs runs a specialized kernel for exchange + DMI in one of the
compact layouts (interfacial along z, x or y, bulk, full blocks) or none +
demag gradient or none, without spin-transfer torque; every other term set
runs the general kernel.  Both take a row's slots in the same order with the
same roundings, so the outputs must be equal bit for bit.  The general kernel
is forced by storing the same DMI components in reverse order (a layout the
fast form does not take): the DMI rows read the same values either way.
Cases: Kuhn box with and without jitter (graph widths 15 and up to 27),
interfacial DMI along z and x, bulk DMI, both together, with and without the
demag field, with uniaxial anisotropy and a uniform external field.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
MU0 = 4e-7 * math.pi


def _mesh(jitter):
    from smx import mesh as M
    m = M.kuhn_box(9, 8, 5, 27.0, 24.0, 15.0, name="box")
    if jitter:
        X = m.X + np.random.default_rng(5).uniform(-jitter, jitter, m.X.shape)
        m = M.TetMesh(X=X, tets=m.tets, name="boxj")
    return m


def _reverse_dmi_storage(hef):
    """Store the DMI components in reverse order; the rows read the same values."""
    off = np.asarray(hef.dmi_off)
    k = int(np.count_nonzero(off >= 0))
    assert k >= 2
    v = hef.dmi.vals.view(-1, k).flip(1).contiguous().view(-1)
    hef.dmi = type(hef.dmi)(nrows=hef.dmi.nrows, k=k, slice_ptr=hef.dmi.slice_ptr, cols=hef.dmi.cols, vals=v,
                            nnz=hef.dmi.nnz)
    new = off.copy()
    new[off >= 0] = k - 1 - off[off >= 0]
    hef.dmi_off = new.astype(np.int32)


def _outputs(hef, m):
    from smx import _capi
    N = hef.N
    rhs = torch.zeros(3 * N, dtype=torch.float64, device=DEV)
    h = torch.zeros(3 * N, dtype=torch.float64, device=DEV)
    if hef.demag_field is not None:
        hef.demag_field.solve_potential(m, speculative=False)
    a = hef._args(m, h_eff=h, rhs=rhs)
    _capi.check(_capi.lib().smx_field_llg_rhs(a, _capi.stream_ptr()), "smx_field_llg_rhs")
    torch.cuda.synchronize()
    out = [rhs.cpu().numpy(), h.cpu().numpy()]
    if hef.demag_field is not None:
        out.append(hef.H_demag_gpu.cpu().numpy())
    return out


@pytest.mark.parametrize("jitter", [0.0, 0.6])
@pytest.mark.parametrize("dmi", ["int_z", "int_x", "bulk", "both"])
@pytest.mark.parametrize("demag", [False, True])
def test_fast_field_kernel_equals_general(jitter, dmi, demag):
    from smx.llg import EffectiveField
    print(f"[field fast] jitter={jitter} dmi={dmi} demag={demag}", flush=True)
    m = _mesh(jitter)
    kw = dict(Ms=8e5, Aex=1.3e-11, alpha=0.3, Ku=5e5, n_ani_vec=(0.2, 0.1, 1.0),
              H0_static=(1e4, -2e4, 3e4), device=DEV)
    if dmi in ("int_z", "both"):
        kw.update(D_int=3e-3, n0_int_vec=(0.0, 0.0, 1.0))
    if dmi == "int_x":
        kw.update(D_int=2e-3, n0_int_vec=(1.0, 0.0, 0.0))
    if dmi in ("bulk", "both"):
        kw.update(D_bulk=1.5e-3)
    if demag:
        kw.update(use_demag=True, demag_kwargs=dict(ksp_rtol_u1=1e-10, ksp_rtol_u2=1e-10))
    hef = EffectiveField(m, **kw)
    g = np.random.default_rng(2).standard_normal((m.num_nodes, 3))
    mv = torch.from_numpy((g / np.linalg.norm(g, axis=1)[:, None]).ravel()).to(DEV)
    fast = _outputs(hef, mv)
    assert np.all(np.isfinite(fast[0])) and np.linalg.norm(fast[0]) > 0.0
    _reverse_dmi_storage(hef)
    general = _outputs(hef, mv)
    for name, f, gen in zip(("rhs", "h_eff", "h_demag"), fast, general):
        bad = np.flatnonzero(f != gen)
        assert bad.size == 0, (name, bad.size, float(np.max(np.abs(f - gen)) / np.max(np.abs(gen))))


@pytest.mark.parametrize("jitter", [0.0, 0.6])
@pytest.mark.parametrize("dmi", ["none", "int_z", "bulk"])
def test_fast_jvp_equals_general(jitter, dmi):
    """smx_llg_jvp's fast form (exchange + DMI layout + uniaxial) against the
    general kernel, bit for bit (general kernel forced as above)."""
    from smx.llg import EffectiveField
    print(f"[jvp fast] jitter={jitter} dmi={dmi}", flush=True)
    m = _mesh(jitter)
    kw = dict(Ms=8e5, Aex=1.3e-11, alpha=0.3, Ku=5e5, n_ani_vec=(0.2, 0.1, 1.0), device=DEV)
    if dmi == "int_z":
        kw.update(D_int=3e-3, n0_int_vec=(0.0, 0.0, 1.0))
    if dmi == "bulk":
        kw.update(D_bulk=1.5e-3)
    hef = EffectiveField(m, **kw)
    rng = np.random.default_rng(4)
    g = rng.standard_normal((m.num_nodes, 3))
    mv = torch.from_numpy((g / np.linalg.norm(g, axis=1)[:, None]).ravel()).to(DEV)
    vv = torch.from_numpy(rng.standard_normal(3 * m.num_nodes)).to(DEV)
    hef.update_jac_state(mv)
    fast = torch.zeros_like(vv)
    hef.jac_vec_times_vec(vv, fast)
    torch.cuda.synchronize()
    assert torch.isfinite(fast).all() and float(fast.norm()) > 0.0
    if dmi == "none":
        return                                    # no layout to force the general kernel with
    _reverse_dmi_storage(hef)
    hef.update_jac_state(mv)
    general = torch.zeros_like(vv)
    hef.jac_vec_times_vec(vv, general)
    torch.cuda.synchronize()
    f, gen = fast.cpu().numpy(), general.cpu().numpy()
    bad = np.flatnonzero(f != gen)
    assert bad.size == 0, (bad.size, float(np.max(np.abs(f - gen)) / np.max(np.abs(gen))))
