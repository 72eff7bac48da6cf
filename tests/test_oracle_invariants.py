"""Pins of the oracle that do not need the reference (CPU).

Each check is chosen so that a plausible mistake in the restated algorithm (a dropped
term, a wrong sign or index, a transposed operand, a wrong dt subset) fails it:
quadrature/basis identities (orthogonality under an independent rule, the divergence
theorem linking derivative, edge and normal tables), free-stream preservation,
conservation (periodic meshes; mass and energy in a box of slip walls), closed forms of the
LF and Roe fluxes on piecewise-constant data, the Roe flux's consistency, exact stationary
contact and shock, and antisymmetry, the reference's dt subset, SSP-RK3 temporal order,
spatial convergence on the exact isentropic vortex, and TVB limiter identities.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from rkdg_heter_calculate_amd import Grid, vortex_params
from rkdg_heter_calculate_amd._lib import BC_OUTFLOW

G = 1.4


def tri_sample(n=40, seed=5):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        x = rng.uniform(-2, 2, 3)
        y = rng.uniform(-2, 2, 3)
        cr = (x[1] - x[0]) * (y[2] - y[0]) - (y[1] - y[0]) * (x[2] - x[0])
        if abs(cr) < 0.2:
            continue
        if cr < 0:
            x[[1, 2]] = x[[2, 1]]
            y[[1, 2]] = y[[2, 1]]
        out.append((x, y))
    return out


@pytest.mark.parametrize("P", [1, 2, 3])
def test_basis_orthogonal_under_volume_rule(P):
    for x, y in tri_sample():
        I = po.triangle_info(P, x, y)
        w, bf = I["vol_w"], I["vol_bf"]
        M = (bf * w) @ bf.T
        assert np.allclose(np.diag(M), I["mass"], rtol=1e-13)
        off = M - np.diag(np.diag(M))
        scale = np.sqrt(np.outer(np.diag(M), np.diag(M)))
        # 15-digit quadrature literals and monomial-form evaluation limit the last digits
        # (absolute off-diagonals stay ~1e-16; the cubic modes have norms down to ~1e-8)
        assert np.abs(off / scale).max() < {1: 1e-12, 2: 1e-11, 3: 1e-8}[P]
        assert abs(I["mass"][0] - I["area"]) < 1e-13 * I["area"]
        assert np.allclose(bf[0], 1.0)
        assert abs(w.sum() - I["area"]) < 1e-14 * I["area"] * 10
        assert abs(I["edge_w"].sum() - I["perimeter"]) < 1e-13 * I["perimeter"]


@pytest.mark.parametrize("P", [1, 2, 3])
def test_divergence_theorem_links_derivative_edge_and_normal_tables(P):
    """sum_q w_q dphi_i/dx = sum_edges sum_j w_j phi_i n_x (both rules exact here)."""
    Nqe = 3 if P <= 2 else 4
    for x, y in tri_sample(20, 9):
        I = po.triangle_info(P, x, y)
        Np = I["mass"].size
        n = I["normal"].reshape(3, 2)
        for d in (0, 1):
            vol = I["vol_bdf"][d::2] @ I["vol_w"]
            nrm = np.repeat(n[:, d], Nqe)
            surf = I["edge_bf"] @ (I["edge_w"] * nrm)
            assert np.allclose(vol, surf, atol=1e-12 * np.abs(surf).max() + 1e-12), (P, d)
        assert Np == (P + 1) * (P + 2) // 2


def test_leading_coefficients_of_the_gram_schmidt_basis():
    for x, y in tri_sample(10, 2):
        I = po.triangle_info(3, x, y)
        A = I["area"]
        c = I["coeff"]
        sq = np.sqrt(A)
        assert c[0, 0] == 1 and np.all(c[0, 1:] == 0)
        assert abs(c[1, 1] - 1 / sq) < 1e-15 * (1 / sq) * 4 and c[1, 2] == 0
        assert abs(c[2, 2] - 1 / sq) < 1e-15 * (1 / sq) * 4
        for k, slot in ((3, 3), (4, 4), (5, 5)):
            assert abs(c[k, slot] - 1 / A) < 1e-14 / A
            assert np.all(c[k, slot + 1:6] == 0)
        for k, slot in ((6, 6), (7, 7), (8, 8), (9, 9)):
            assert abs(c[k, slot] - 1 / (A * sq)) < 1e-13 / (A * sq)
            assert np.all(c[k, slot + 1:] == 0)


def periodic_grid(n=8, jitter=0.2, seed=1):
    return Grid.splitquad(n, n, 0.0, 0.0, 10.0, 10.0, periodic=True, diagonal=1, jitter=jitter, seed=seed)


def oracle_for(g, **kw):
    ex, ey = g.element_coords()
    return po.Oracle(ex, ey, g.neighbour, g.shared_edge, g.boundary, **kw)


@pytest.mark.parametrize("P", [1, 2, 3])
@pytest.mark.parametrize("flux", [0, 1])
def test_free_stream_preserved(P, flux):
    g = periodic_grid()
    o = oracle_for(g, P=P, flux=flux)
    U = np.zeros(o.shape())
    U[:, 0, :] = np.array([1.2, 0.3, -0.4, 3.1])[:, None]
    o.set_state(U)
    R = o.residual()
    assert np.abs(R).max() < 1e-12


@pytest.mark.parametrize("flux", [0, 1])
def test_conservation_on_periodic_mesh(flux):
    g = periodic_grid(10)
    o = oracle_for(g, P=2, flux=flux)
    o.project(1, vortex_params())
    R = o.residual()
    tot = R[:, 0, :].sum(axis=1)
    assert np.abs(tot).max() < 1e-11 * np.abs(R[:, 0, :]).sum(axis=1).max()


def euler_F(U):
    r, mu, mv, E = U
    u, v = mu / r, mv / r
    p = (G - 1) * (E - 0.5 * r * (u * u + v * v))
    return np.array([mu, mu * u + p, mu * v, u * (E + p)]), np.array([mv, mv * u, mv * v + p, v * (E + p)])


def piecewise_constant(g, P, seed):
    rng = np.random.default_rng(seed)
    E = g.n_elements
    Np = (P + 1) * (P + 2) // 2
    U = np.zeros((4, Np, E))
    r = rng.uniform(0.8, 1.2, E)
    u = rng.uniform(-0.5, 0.5, E)
    v = rng.uniform(-0.5, 0.5, E)
    p = rng.uniform(0.8, 1.2, E)
    U[0, 0], U[1, 0], U[2, 0], U[3, 0] = r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)
    return U


@pytest.mark.parametrize("flux", [0, 1])
def test_surface_flux_closed_form_on_piecewise_constant_data(flux):
    """Mode-0 residual of piecewise-constant data is -sum_k l_k Fhat_k (volume term vanishes).
    LF: alpha from the arithmetic mean of the two cell means (src/kernelfunctions.cu:378-386)."""
    g = periodic_grid(6, seed=4)
    o = oracle_for(g, P=1, flux=flux)
    U = piecewise_constant(g, 1, 7)
    o.set_state(U)
    R = o.residual()
    ex, ey = g.element_coords()
    for e in range(g.n_elements):
        acc = np.zeros(4)
        for k in range(3):
            a, b = (k + 1) % 3, (k + 2) % 3
            l = np.hypot(ex[e, a] - ex[e, b], ey[e, a] - ey[e, b])
            nx, ny = (ey[e, b] - ey[e, a]) / l, (ex[e, a] - ex[e, b]) / l
            n = g.neighbour[e, k]
            UL, UR = U[:, 0, e], U[:, 0, n]
            if flux == 0:
                m = 0.5 * (UL + UR)
                uu, vv = m[1] / m[0], m[2] / m[0]
                pp = (G - 1) * (m[3] - 0.5 * m[0] * (uu * uu + vv * vv))
                alpha = abs(uu * nx + vv * ny) + np.sqrt(G * pp / m[0])
                FL, GL = euler_F(UL)
                FR, GR = euler_F(UR)
                Fh = 0.5 * ((FL + FR) * nx + (GL + GR) * ny - alpha * (UR - UL))
            else:
                Fh = po.roe_flux(G, UL, UR, nx, ny)
            acc -= l * Fh
        assert np.allclose(R[:, 0, e], acc, rtol=1e-12, atol=1e-13), e


def test_roe_properties():
    rng = np.random.default_rng(0)
    for _ in range(20):
        r, u, v, p = rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2)
        U = np.array([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)])
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(th), np.sin(th)
        F, Gf = euler_F(U)
        assert np.allclose(po.roe_flux(G, U, U, nx, ny), F * nx + Gf * ny, rtol=1e-14, atol=1e-14)
    # stationary contact: equal pressure, zero normal velocity, different density
    nx, ny = 0.6, 0.8
    t = 0.3  # tangential velocity
    UL = np.array([1.0, -t * ny * 1.0, t * nx * 1.0, 1.0 / (G - 1) + 0.5 * 1.0 * t * t])
    UR = np.array([0.25, -t * ny * 0.25, t * nx * 0.25, 1.0 / (G - 1) + 0.5 * 0.25 * t * t])
    Fh = po.roe_flux(G, UL, UR, nx, ny)
    assert np.allclose(Fh, [0, 1.0 * nx, 1.0 * ny, 0], atol=1e-14)
    # stationary normal shock (Rankine-Hugoniot, Mach 2 upstream)
    M1 = 2.0
    r1, p1 = 1.0, 1.0
    u1 = M1 * np.sqrt(G * p1 / r1)
    r2 = r1 * (G + 1) * M1 ** 2 / ((G - 1) * M1 ** 2 + 2)
    p2 = p1 * (2 * G * M1 ** 2 - (G - 1)) / (G + 1)
    u2 = r1 * u1 / r2
    UL = np.array([r1, r1 * u1 * nx, r1 * u1 * ny, p1 / (G - 1) + 0.5 * r1 * u1 ** 2])
    UR = np.array([r2, r2 * u2 * nx, r2 * u2 * ny, p2 / (G - 1) + 0.5 * r2 * u2 ** 2])
    F, Gf = euler_F(UL)
    assert np.allclose(po.roe_flux(G, UL, UR, nx, ny), F * nx + Gf * ny, rtol=1e-12, atol=1e-12)
    # supersonic flow to the right: pure upwinding
    UR2 = UL * 1.1
    assert np.allclose(po.roe_flux(G, UL, UR2, nx, ny), F * nx + Gf * ny, rtol=1e-12, atol=1e-12)


def test_dt_subset_of_the_reference_reduction():
    """dt = cfl / max over cells with id % 128 < 64 (kernel_getTimeStep's reduction)."""
    g = periodic_grid(16, seed=8)  # 512 elements
    U = piecewise_constant(g, 1, 3)
    E = g.n_elements
    # plant the fastest cell in the dropped half
    U[1, 0, 100] = U[0, 0, 100] * 5.0
    U[3, 0, 100] += 0.5 * (U[1, 0, 100] ** 2) / U[0, 0, 100]
    ex, ey = g.element_coords()
    area = 0.5 * np.abs((ex[:, 2] - ex[:, 0]) * (ey[:, 1] - ey[:, 0]) - (ey[:, 2] - ey[:, 0]) * (ex[:, 1] - ex[:, 0]))
    per = sum(np.hypot(ex[:, (k + 1) % 3] - ex[:, (k + 2) % 3], ey[:, (k + 1) % 3] - ey[:, (k + 2) % 3])
              for k in range(3))
    r, u, v = U[0, 0], U[1, 0] / U[0, 0], U[2, 0] / U[0, 0]
    p = (G - 1) * (U[3, 0] - 0.5 * r * (u * u + v * v))
    s = (np.sqrt(u * u + v * v) + np.sqrt(G * p / r)) * per / area
    keep = (np.arange(E) % 128) < 64
    for pol, R, expect in ((0, 512, 0.3 / s[keep].max()), (1, 512, 0.3 / s.max()), (0, 64, 0.3 / s.max())):
        o = oracle_for(g, P=1, dt_policy=pol, cfl=0.3, reduction_threads=R)
        o.set_state(U)
        assert abs(o.compute_dt() - expect) < 1e-14 * expect, (pol, R)
    assert s.argmax() == 100 and not keep[100]


def test_ssp_rk3_third_order_in_time():
    g = periodic_grid(8, jitter=0.1, seed=3)
    T = 0.4
    res = {}
    for n in (8, 16, 32, 64):
        o = oracle_for(g, P=2, dt_policy=2, dt_fixed=T / n)
        o.project(1, vortex_params())
        o.advance(n)
        res[n] = o.get_state()
    M = oracle_for(g, P=2).mass()
    e1 = po.l2_norms(res[8] - res[64], M).max()
    e2 = po.l2_norms(res[16] - res[64], M).max()
    e3 = po.l2_norms(res[32] - res[64], M).max()
    r1, r2 = np.log2(e1 / e2), np.log2(e2 / e3)
    # differences against the finest run: ratio -> 8 as the finest error vanishes
    assert r1 > 2.7 and r2 > 2.6, (r1, r2)


@pytest.mark.parametrize("P,order", [(1, 1.9), (2, 2.3), (3, 3.6)])
def test_vortex_convergence_order(P, order):
    """Error against the projected exact vortex at T = 0.2 on 16/32/64 meshes (LF gives
    about 2, 2.5 and 4 for P1, P2, P3 on these split-quad meshes)."""
    errs = []
    T = 0.2
    for n in (16, 32, 64):
        g = Grid.splitquad(n, n, 0.0, 0.0, 10.0, 10.0, periodic=True, diagonal=0)
        o = oracle_for(g, P=P, dt_policy=2, dt_fixed=T / (2 * n))
        o.project(1, vortex_params())
        o.advance(2 * n)
        U = o.get_state()
        ex = oracle_for(g, P=P)
        ex.project(1, vortex_params(t=T))
        M = ex.mass()
        errs.append(po.l2_norms(U - ex.get_state(), M)[0])
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > order, (errs, rates)


def test_tvb_limiter_identities():
    g = Grid.splitquad(12, 12, 0.0, 0.0, 1.0, 1.0, periodic=False, diagonal=1, jitter=0.2, seed=6, bc=BC_OUTFLOW)
    ex, ey = g.element_coords()
    E = g.n_elements
    # linear data: untouched away from the boundary, means preserved everywhere
    o = oracle_for(g, P=1, limiter=1, tvb_M=0.0, dt_policy=2, dt_fixed=1.0)
    lin = po.Oracle(ex, ey, g.neighbour, g.shared_edge, g.boundary, P=1)
    Ul = np.zeros((4, 3, E))
    # project u = 1 + 0.3 x - 0.2 y in each variable via the 3-mode fit at the centroid
    for e in range(E):
        I = po.triangle_info(1, ex[e], ey[e])
        bx, by = I["bary"]
        # phi1 = X, phi2 = Y + d X + f (scaled): solve for the modes of a linear function
        c = I["coeff"]
        A = np.array([[c[1, 1], c[2, 1]], [c[1, 2], c[2, 2]]])
        a1, a2 = np.linalg.solve(A, [0.3, -0.2])
        mean = 1 + 0.3 * bx - 0.2 * by
        Ul[:, 0, e] = mean - a2 * c[2, 0]
        Ul[:, 1, e] = a1
        Ul[:, 2, e] = a2
        Ul[3, 0, e] += 2.0  # keep pressure positive for the state's other uses
    del lin
    o.set_state(Ul)
    o.stage(0, 0.0)  # dt = 0: identity update, then the limiter
    Uo = o.get_state()
    f = o.flags()
    interior = (g.neighbour >= 0).all(axis=1)
    # interior cells whose neighbours are also interior see exact linear data
    nb_interior = np.array([interior[e] and interior[g.neighbour[e]].all() for e in range(E)])
    assert np.array_equal(Uo[:, 0], Ul[:, 0])
    assert np.allclose(Uo[:, :, nb_interior], Ul[:, :, nb_interior], rtol=1e-10, atol=1e-12)
    assert f[nb_interior].sum() == 0
    # rough data: means preserved, some cells troubled; M -> infinity changes nothing
    rng = np.random.default_rng(1)
    Ur = Ul.copy()
    Ur[:, 1:, :] += rng.normal(0, 0.2, Ur[:, 1:, :].shape)
    o.set_state(Ur)
    o.stage(0, 0.0)
    assert np.array_equal(o.get_state()[:, 0], Ur[:, 0])
    assert o.flags().sum() > 0
    o2 = oracle_for(g, P=1, limiter=1, tvb_M=1e30, dt_policy=2, dt_fixed=1.0)
    o2.set_state(Ur)
    o2.stage(0, 0.0)
    assert np.array_equal(o2.get_state(), Ur) and o2.flags().sum() == 0


def test_positivity_scaling_keeps_point_pressure_positive():
    """After the limiter pass, density and pressure at every volume and edge point are
    positive, also for data whose unlimited traces are not (P1, low-pressure state)."""
    g = Grid.splitquad(10, 10, 0.0, 0.0, 1.0, 1.0, periodic=False, diagonal=1, jitter=0.2, seed=3, bc=BC_OUTFLOW)
    E = g.n_elements
    ex, ey = g.element_coords()
    rng = np.random.default_rng(5)
    U = np.zeros((4, 3, E))
    r = rng.uniform(0.1, 0.2, E)
    u = rng.uniform(1.0, 1.3, E)
    v = rng.uniform(1.0, 1.3, E)
    p = rng.uniform(0.02, 0.04, E)
    U[0, 0], U[1, 0], U[2, 0], U[3, 0] = r, r * u, r * v, p / 0.4 + 0.5 * r * (u * u + v * v)
    U[:, 1:, :] = rng.normal(0, 1, (4, 2, E)) * np.abs(U[:, :1, :]) * 0.8
    o = oracle_for(g, P=1, limiter=1, tvb_M=1e30, dt_policy=2, dt_fixed=1.0)  # TVB inactive
    o.set_state(U)
    o.stage(0, 0.0)
    V = o.get_state()
    assert np.array_equal(V[:, 0], U[:, 0])
    worst_before, worst_after = np.inf, np.inf
    for e in range(E):
        I = po.triangle_info(1, ex[e], ey[e])
        bf = np.concatenate([I["vol_bf"], I["edge_bf"]], axis=1)
        for W, name in ((U, "before"), (V, "after")):
            pts = np.einsum("vi,iq->vq", W[:, :, e], bf)
            pr = 0.4 * (pts[3] - 0.5 * (pts[1] ** 2 + pts[2] ** 2) / pts[0])
            m = min(pts[0].min(), pr.min()) if pts[0].min() > 0 else -1.0
            if name == "before":
                worst_before = min(worst_before, m)
            else:
                worst_after = min(worst_after, m)
    assert worst_before < 0 < worst_after
    assert (o.flags() & 2).sum() > 0 and (o.flags() & 1).sum() == 0


def test_roe_flux_is_antisymmetric():
    """F^(UR, UL, -n) = -F^(UL, UR, n), entropy fix included (the fix's wave-speed bounds map
    onto each other when the states swap and the normal turns): each interior edge's two
    element-owned evaluations agree."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        def state():
            r, u, v, p = rng.uniform(0.2, 3), rng.uniform(-2, 2), rng.uniform(-2, 2), rng.uniform(0.2, 3)
            return np.array([r, r * u, r * v, p / (G - 1) + 0.5 * r * (u * u + v * v)])
        UL, UR = state(), state()
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(th), np.sin(th)
        a = po.roe_flux(G, UL, UR, nx, ny)
        b = po.roe_flux(G, UR, UL, -nx, -ny)
        assert np.allclose(a, -b, rtol=1e-12, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("flux", [0, 1])
def test_reflecting_box_conserves_mass_and_energy(flux):
    """A closed box of slip walls: the mirrored ghost state (reflection about the edge normal)
    gives zero mass and energy flux through every wall edge, so the mass and energy residuals
    sum to zero; the momentum does not (the walls' pressure). A wrong sign in the mirrored
    velocity makes mass leak."""
    from rkdg_heter_calculate_amd._lib import BC_REFLECTION
    g = Grid.splitquad(10, 10, 0.0, 0.0, 10.0, 10.0, periodic=False, diagonal=1, jitter=0.2, seed=3, bc=BC_REFLECTION)
    o = oracle_for(g, P=2, flux=flux)
    o.project(1, vortex_params(uinf=0.4, vinf=-0.3, xc=5.0, yc=5.0, lx=0.0, ly=0.0))
    R = o.residual()
    tot = R[:, 0, :].sum(axis=1)
    scale = np.abs(R[:, 0, :]).sum(axis=1)
    assert abs(tot[0]) < 1e-12 * scale[0], tot
    assert abs(tot[3]) < 1e-12 * scale[3], tot
    assert max(abs(tot[1]) / scale[1], abs(tot[2]) / scale[2]) > 1e-3, tot
