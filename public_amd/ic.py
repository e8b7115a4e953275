"""Seeded initial conditions and grid set-up for the periodic-box workloads.

The reference reads its initial state through a case-level
``initial_condition`` module; this module produces the synthetic states the
workloads are defined on:

* a solenoidal random velocity field with energy spectrum
  E(k) ~ k^4 exp(-2 (k/k0)^2), k0 = 4, random Fourier phases drawn from the
  seed, projected divergence-free and scaled so that the per-component rms
  velocity sqrt(<u^2+v^2+w^2>/3) is 1;
* rho = 1 and p = 1/(gamma Ma^2), so that the rms Mach number is Ma;
* dt from CFL 0.5 using the reference's own definition
  cfl = max(|u| + c) dt / min(dx)  (src/helper/check_state.f90:125-136),
  evaluated once on the initial state and then kept fixed.

State layout is the reference's: q(n1, n2, n3, 5) Fortran order,
variables (rho, rho*u, rho*v, rho*w, p).
"""
import numpy as np


def periodic_dx(n, length=2.0 * np.pi):
    """dx = L / n for a periodic core (src/geometry/grid_transform.f90:41-57)."""
    return length / n


def solenoidal_velocity(n, seed, k0=4.0):
    """Return (u, v, w), each (n, n, n) Fortran-ordered, on [0, 2pi)^3."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kx, ky, kz = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    spec = kk ** 4 * np.exp(-2.0 * (kk / k0) ** 2)
    # amplitude per mode so that the shell-summed energy follows E(k)
    amp = np.zeros_like(kk)
    nz = kk > 0
    amp[nz] = np.sqrt(spec[nz] / (4.0 * np.pi * kk[nz] ** 2))
    uh = []
    for _ in range(3):
        phase = rng.uniform(0.0, 2.0 * np.pi, size=kk.shape)
        mag = rng.standard_normal(size=kk.shape)
        uh.append(amp * mag * np.exp(1j * phase))
    # project onto the divergence-free subspace: u_hat -= k (k . u_hat)/|k|^2
    kd = np.where(nz, kk * kk, 1.0)
    dot = (kx * uh[0] + ky * uh[1] + kz * uh[2]) / kd
    uh = [uh[0] - kx * dot, uh[1] - ky * dot, uh[2] - kz * dot]
    # Nyquist planes have no conjugate partner: drop them so the field is real and solenoidal
    if n % 2 == 0:
        ny = n // 2
        for c in uh:
            c[ny, :, :] = 0
            c[:, ny, :] = 0
            c[:, :, ny] = 0
    vel = [np.real(np.fft.ifftn(c)) for c in uh]
    rms = np.sqrt((vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2).mean() / 3.0)
    return [np.asfortranarray(c / rms) for c in vel]


def periodic_box_state(n, seed, mach, gamma=1.4):
    """q(n, n, n, 5) for the periodic box at rms Mach number `mach`."""
    u, v, w = solenoidal_velocity(n, seed)
    q = np.zeros((n, n, n, 5), order="F")
    q[..., 0] = 1.0
    q[..., 1] = u
    q[..., 2] = v
    q[..., 3] = w
    q[..., 4] = 1.0 / (gamma * mach * mach)
    return q


def cfl_dt(q, dx_min, gamma=1.4, cfl=0.5):
    """dt such that max(|u|+c) dt / dx_min = cfl (check_state.f90:125-136)."""
    rho = q[..., 0]
    u = q[..., 1] / rho
    v = q[..., 2] / rho
    w = q[..., 3] / rho
    c = np.sqrt(gamma * q[..., 4] / rho)
    umax = float(np.max(np.sqrt(u * u + v * v + w * w) + c))
    return cfl * dx_min / umax


def rho_E(q, gamma=1.4):
    """Total energy rho*E = p/(gamma-1) + rho |u|^2 / 2 from the (rho, rho u, p) state."""
    rho = q[..., 0]
    ke = 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / rho
    return q[..., 4] / (gamma - 1.0) + ke


def kinetic_energy(q):
    """Volume-mean kinetic energy 0.5 <rho |u|^2>."""
    rho = q[..., 0]
    return float((0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / rho).mean())


def bandlimited_velocity(n, seed, k0=4.0, kmax=16):
    """(u, v, w) of the same spectrum, drawn only for wavenumbers in
    [-kmax, kmax)^3 and summed on the grid by separable partial DFTs.

    E(k) ~ k^4 exp(-2 (k/k0)^2) is below 1e-10 of its peak beyond k = 16, so
    the truncation changes nothing measurable, and the field is the same
    function on every grid finer than 2 kmax: a 256^3 state costs three
    matrix products instead of full FFTs, and a coarse grid can be given
    the very same field.  Same seed, other draws than `solenoidal_velocity`."""
    rng = np.random.default_rng(seed)
    k1 = np.arange(-kmax, kmax, dtype=np.float64)
    kx, ky, kz = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    nz = kk > 0
    amp = np.zeros_like(kk)
    amp[nz] = np.sqrt(kk[nz] ** 4 * np.exp(-2.0 * (kk[nz] / k0) ** 2) / (4.0 * np.pi * kk[nz] ** 2))
    uh = []
    for _ in range(3):
        phase = rng.uniform(0.0, 2.0 * np.pi, size=kk.shape)
        mag = rng.standard_normal(size=kk.shape)
        uh.append(amp * mag * np.exp(1j * phase))
    kd = np.where(nz, kk * kk, 1.0)
    dot = (kx * uh[0] + ky * uh[1] + kz * uh[2]) / kd
    uh = [uh[0] - kx * dot, uh[1] - ky * dot, uh[2] - kz * dot]
    x = 2.0 * np.pi * np.arange(n) / n
    E = np.exp(1j * np.outer(x, k1))  # (n, 2 kmax)
    vel = []
    for c in uh:
        a = c @ E.T                                   # (kx, ky, z)
        a = np.einsum("ybz,jb->yjz", a, E)            # (kx, y, z)
        a = (E @ a.reshape(a.shape[0], -1)).reshape(n, n, n)  # (x, y, z)
        vel.append(np.real(a))
    rms = np.sqrt((vel[0] ** 2 + vel[1] ** 2 + vel[2] ** 2).mean() / 3.0)
    return [np.asfortranarray(c / rms) for c in vel]


def bandlimited_box_state(n, seed, mach, gamma=1.4, kmax=16):
    """q(n, n, n, 5) as `periodic_box_state` from `bandlimited_velocity`."""
    u, v, w = bandlimited_velocity(n, seed, kmax=kmax)
    q = np.zeros((n, n, n, 5), order="F")
    q[..., 0] = 1.0
    q[..., 1] = u
    q[..., 2] = v
    q[..., 3] = w
    q[..., 4] = 1.0 / (gamma * mach * mach)
    return q
