"""GPU face-loop geometry and assembly against the oracle: bit-identical.

The oracle is itself bit-identical to the reference's build_fv_geometry,
gauss_gradient_with_boundary, make_mass_flux, the SIMPLE corrector assembly
and assemble_cell_diffusion_matrix (test_oracle_reference.py), so equality
here is equality with the reference.
"""
import numpy as np
import pytest

import cfdx_dev_amd as cx
from systems import box_mesh

pytestmark = pytest.mark.gpu

MESHES = [
    dict(kind="box", nx=7, ny=6, nz=5, jitter=0.15, permute=True, seed=2),
    dict(kind="box", nx=12, ny=16, nz=8, jitter=0.10, permute=True, seed=5, wall_layers=4, first_aspect=50.0),
    dict(kind="extruded2d", nx=16, ny=16, nz=1, jitter=0.0, permute=False, seed=0),
    dict(kind="channel_step", nx=24, ny=12, nz=4, lx=3.0, jitter=0.05, permute=True, seed=9,
         first_aspect=20.0, growth=1.15, step_height_fraction=1.0 / 3.0, step_length_fraction=0.25),
]


def _mesh(spec):
    spec = dict(spec)
    return cx.generate_mesh(**spec)


@pytest.mark.parametrize("spec", MESHES)
def test_geometry_bit_identical(oracle, spec):
    m = _mesh(spec)
    fv = cx.FvMesh(m)
    g_gpu = fv.geometry()
    g_ora = oracle.geometry(m)
    for a, b in zip(g_gpu, g_ora):
        assert np.array_equal(a, b), np.abs(a - b).max()


@pytest.mark.parametrize("spec", MESHES)
@pytest.mark.parametrize("fixed", [False, True])
def test_pressure_assembly_bit_identical(oracle, spec, fixed):
    m = _mesh(spec)
    fv = cx.FvMesh(m)
    g = oracle.geometry(m)
    rng = np.random.default_rng(11)
    nc = m.n_cells
    rau = np.exp(rng.normal(size=3 * nc))
    U = rng.normal(size=3 * nc)
    p = rng.normal(size=nc)
    pt = np.zeros(m.n_patches, np.uint8)
    pv = np.zeros(m.n_patches)
    vt = np.zeros(m.n_patches, np.uint8)
    vv = np.zeros(3 * m.n_patches)
    vt[0] = 1
    vv[0:3] = [1.0, 0.1, -0.2]
    if fixed:
        pt[1] = 1
        pv[1] = 0.25
        pt[2] = 2  # fixed gradient (zero) on y-min for the gradient operator
    grad_o = oracle.gauss_gradient(m, g, pt, pv, p)
    grad_g = fv.gauss_gradient(pt, pv, p)
    assert np.array_equal(grad_o, grad_g)
    phi_o = oracle.mass_flux(m, g, vt, vv, 1.2, U)
    phi_g = fv.mass_flux(vt, vv, 1.2, U)
    assert np.array_equal(phi_o, phi_g)
    pt_asm = np.where(pt == 1, 1, 0).astype(np.uint8)
    grad_a = oracle.gauss_gradient(m, g, pt_asm, pv, p)
    a = oracle.pressure_assemble(m, g, pt_asm, pv, 1.2, rau, phi_o, p, grad_a, 0)
    rp, cols, vals, rhs = fv.pressure_assemble(pt_asm, pv, 1.2, rau, phi_o, p, 0)
    assert np.array_equal(rp, a["rowp"]) and np.array_equal(cols, a["cols"])
    assert np.array_equal(vals, a["vals"])
    assert np.array_equal(rhs, a["rhs"])


@pytest.mark.parametrize("spec", MESHES[:3])
def test_diffusion_assembly_bit_identical(oracle, spec):
    m = _mesh(spec)
    fv = cx.FvMesh(m)
    g = oracle.geometry(m)
    rng = np.random.default_rng(4)
    fval = np.full(m.n_faces, np.nan)
    nb = m.n_faces - m.n_internal_faces
    fval[m.n_internal_faces:] = rng.normal(size=nb)
    fneu = np.zeros(m.n_faces, np.uint8)
    fneu[m.n_internal_faces::3] = 1
    # EMPTY patches carry no prescribed value
    for f in range(m.n_internal_faces, m.n_faces):
        if m.patch_kind[m.patch_of_face[f]] == 1:
            fval[f] = np.nan
    src = rng.normal(size=m.n_cells)
    o = oracle.diffusion_assemble(m, g, 1.3, src, fval, fneu)
    d = fv.diffusion_assemble(1.3, src, fval, fneu)
    for x, y in zip(o, d):
        assert np.array_equal(x, y)
