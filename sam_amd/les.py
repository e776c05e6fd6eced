"""LES with the learned closure: training data from a DNS, training, the LES
run with its a-posteriori check, and spectral comparisons.

All compute runs in libsamhip (DNS/LES steps, box filter, BF16 MFMA
training and inference); this module only sequences the calls.
"""
import numpy as np

from . import ic
from . import samhip as S


def spectral_truncate(u, n_to, n_dealias=1):
    """Keep the Fourier modes of u (SAM's stored form) that a n_to^3 grid
    holds, as remesh.f:97-178 re-meshes restart files."""
    nzm_f, nym_f, nxm_f, lu = u.shape
    kf = (nym_f - 1) // 2
    d = ic.grid_dims(n_to, n_dealias=n_dealias, i_strat=1 if lu == 4 else 0)
    out = np.zeros((d["nzm"], d["nym"], d["nxm"], lu), np.complex128, order="F")
    ky = np.array([j if j <= d["kym"] else j - d["nym"] for j in range(d["nym"])])
    kz = np.array([k if k <= d["kzm"] else k - d["nzm"] for k in range(d["nzm"])])
    if d["kym"] > kf:
        raise ValueError("target grid is finer than the source")
    jf = ky % nym_f
    kf_idx = kz % nzm_f
    out[:, :, :, :] = u[np.ix_(kf_idx, jf, np.arange(d["nxm"]), np.arange(lu))]
    return out


def shell_energy(spec):
    """0.5 * sum of |u_k|^2 per integer shell (spectra.f's ek / 2)"""
    return 0.5 * np.asarray(spec["ek"])


def collect_samples(dns, mlp, ratio, nsnap, steps_between):
    """advance the DNS handle and add one box-filtered snapshot every
    `steps_between` steps"""
    for _ in range(nsnap):
        dns.step(steps_between)
        mlp.add_samples(dns, ratio)


def train(mlp, nsteps, batch, lr=1e-3, seed=0):
    mlp.fit_norms()
    return mlp.train(nsteps, batch=batch, lr=lr, seed=seed)


def run_les(les, nsteps, check_every=0, tol=1.0):
    """advance an LES handle; with a learned closure attached, check it
    against the dynamic model every `check_every` steps and fall back when
    the SGS transfer spectra disagree by more than `tol` (relative L2)."""
    checks = []
    done = 0
    while done < nsteps:
        chunk = nsteps - done if not check_every else min(check_every, nsteps - done)
        les.step(chunk)
        done += chunk
        if check_every and les.params.sgs == S.SGS_MLP:
            err, switched = S.sgs_check(les, tol)
            checks.append((done, err, switched))
            if switched:
                les.params.sgs = S.SGS_DYN
    return checks
