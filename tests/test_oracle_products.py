"""The slice view restatement (js/main.js:925-965) against closed forms."""
import numpy as np

from oracle.oracle import slice_view

DIMS = (20, 12, 10)


def planes(ux, uy, uz, rho):
    nx, ny, nz = DIMS
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    F = {"ux": ux(x, y, z), "uy": uy(x, y, z), "uz": uz(x, y, z), "rho": rho(x, y, z)}
    return {k: np.asarray(v, dtype=np.float64).reshape(-1) for k, v in F.items()}


def test_solid_body_rotation_has_twice_its_rate_as_vorticity():
    om, xc, yc = 0.003, 9.0, 5.5
    F = planes(lambda x, y, z: -om * (y - yc), lambda x, y, z: om * (x - xc), lambda x, y, z: 0 * x,
               lambda x, y, z: 1 + 0 * x)
    solid = np.zeros(np.prod(DIMS), np.uint8)
    w = slice_view(F, solid, DIMS, "xy", 4, "vorticity", 0.05)
    np.testing.assert_allclose(w[1:-1, 1:-1], 2 * om, rtol=1e-6)
    # one-sided at the edges: half of one difference plus the full other
    np.testing.assert_allclose(w[1:-1, 0], 1.5 * om, rtol=1e-6)
    sp = slice_view(F, solid, DIMS, "xy", 4, "speed", 0.05)
    y, x = np.mgrid[0:DIMS[1], 0:DIMS[0]]
    np.testing.assert_allclose(sp, om * np.hypot(x - xc, y - yc) / 0.05, rtol=1e-6)


def test_xz_vorticity_sign_and_pressure_of_a_linear_density():
    # u_x = a z, u_z = b x: in the xz view w = du_x/dz - du_z/dx = a - b
    a, b = 0.002, -0.001
    F = planes(lambda x, y, z: a * z, lambda x, y, z: 0 * x, lambda x, y, z: b * x,
               lambda x, y, z: 1 + 1e-3 * x + 2e-3 * z)
    solid = np.zeros(np.prod(DIMS), np.uint8)
    w = slice_view(F, solid, DIMS, "xz", 3, "vorticity", 0.05)
    np.testing.assert_allclose(w[1:-1, 1:-1], a - b, rtol=1e-6)
    cp = slice_view(F, solid, DIMS, "xz", 3, "pressure", 0.05)
    z, x = np.mgrid[0:DIMS[2], 0:DIMS[0]]
    rho = 1 + 1e-3 * x + 2e-3 * z
    np.testing.assert_allclose(cp[1:-1, 1:-1], (((rho - 1) / 3) / (0.5 * 0.05 ** 2))[1:-1, 1:-1], rtol=1e-5)


def test_solid_cells_read_zero():
    F = planes(lambda x, y, z: 0.05 + 0 * x, lambda x, y, z: 0 * x, lambda x, y, z: 0 * x, lambda x, y, z: 1 + 0 * x)
    solid = np.zeros(DIMS[::-1], np.uint8)
    solid[4, 5, 7] = 1
    v = slice_view(F, solid.reshape(-1), DIMS, "xy", 4, "speed", 0.05)
    assert v[5, 7] == 0 and v[5, 6] == 1
