"""The packer on the GPU.

* One iteration (s0 = 1, one final iteration) against the oracle's
  restatement on the same input: positions within 1e-12 mm (float64 sums in
  another order, float32 displacements).
* A full packing of a D12 x H16 mm cylinder at solid fraction 0.6: the largest
  overlap is below the tolerance (5 % of the smaller radius) and the mean below
  1.5 %, every sphere lies inside the container, the solid fraction is as
  asked, most spheres touch several others, and a second run with the same
  seed gives the same packing bit for bit.
"""
import numpy as np
import pytest

from oracle.oracle import pack_iteration
from rfd import native, native_cases

pytestmark = pytest.mark.gpu


def test_one_iteration_matches_oracle(ctx):
    g = np.random.default_rng(8)
    N, Rc, H = 300, 3.0, 6.0
    rad = g.uniform(0.4, 0.6, N)
    rho = (Rc - 0.6) * np.sqrt(g.uniform(0, 1, N))
    th = g.uniform(0, 2 * np.pi, N)
    R = np.stack([rho * np.cos(th), rho * np.sin(th), g.uniform(-H / 2 + 0.6, H / 2 - 0.6, N)], 1)
    p = native.params_from_material(native_cases.material_c(0.4), refmode=0,
                                    contact_capacity=64)
    s = native.NativeSystem(ctx, p, R, rad)
    s.set_bonds(native.Bonds(*(np.zeros(0, t) for t in (np.uint32, np.uint32, np.float32,
                                                        np.float32, np.float32))))
    pp = native.PackParams(cx=0.0, cy=0.0, radius=Rc, z0=-H / 2, z1=H / 2, s0=1.0,
                           growth=1.01, beta=0.5, momentum=0.0, tol=1e-30, stage_iters=1,
                           final_iters=1)
    res = s.pack(pp)
    assert res.iterations == 1 and not res.row_overflow
    got = s.state()["R"].reshape(3, N).T
    # the device keeps radii in FP32 (set_particles), the oracle takes the same
    want, mo = pack_iteration(R, rad.astype(np.float32), 0.5, 0.0, 0.0, Rc, -H / 2, H / 2)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    assert res.max_overlap == pytest.approx(mo, rel=1e-6)
    s.close()


def test_full_packing(ctx):
    D, H, phi = 12.0, 16.0, 0.6
    a = native_cases.packed_cylinder(ctx, D, H, phi, seed=5)
    b = native_cases.packed_cylinder(ctx, D, H, phi, seed=5)
    print("packing", len(a.radius), a.result, flush=True)
    assert a.result["converged"] and not a.result["row_overflow"]
    np.testing.assert_array_equal(a.R, b.R)
    R, r = a.R, a.radius
    assert abs(a.phi - phi) < 0.01
    rho = np.hypot(R[:, 0], R[:, 1])
    assert np.all(rho + r <= D / 2 + 1e-9)
    assert np.all(R[:, 2] - r >= -H / 2 - 1e-9) and np.all(R[:, 2] + r <= H / 2 + 1e-9)
    from scipy.spatial import cKDTree

    P = cKDTree(R).query_pairs(1.2, output_type="ndarray")
    d = np.linalg.norm(R[P[:, 1]] - R[P[:, 0]], axis=1)
    ov = (r[P[:, 0]] + r[P[:, 1]] - d) / np.minimum(r[P[:, 0]], r[P[:, 1]])
    assert ov.max() < 0.05
    assert ov[ov > 0].mean() < 0.015
    # a packing, not a gas: most spheres touch (gap < 0.05 r_mean) a few others
    touch = np.bincount(P[(r[P[:, 0]] + r[P[:, 1]] - d) > -0.025].ravel(), minlength=len(r))
    assert touch.mean() > 2.0
