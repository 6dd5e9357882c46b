"""Training and serving of the emulator on spectra the GPU batch generates (the fifth workload).

    data = generate(dctx, n, seed)               # parameters -> spectra and magnitudes by opensps_hip_csp_batch
    pca = fit_pca(data_spec_sample, 128)         # the PCA basis of the log-spectra
    emu = Emulator.train(dctx, cosmo, ...)       # targets [PCA coefficients, magnitudes], Adam with cosine decay
    out = emu.predict(params)                    # the network, with the batch as fallback and judge

Inputs (an InputSpace): the per-model parameters the parametric workload varies (tage, tau, logzsol, dust2,
dust_index, zred, sigma_smooth), uniform over its ranges (BASE); or those and the three IMF slopes (the batch on the
SSP operator) and the dust emission and AGN parameters (duste_umin, duste_qpah, duste_gamma, fagn; the context holds
the templates) (FULL, 14 inputs). The log-spectrum is log10(max(spec, 1e-20 x the model's peak)) on the grid;
the magnitudes are clipped at the model's brightest band + 30 mag (the faintest bands at high redshift, below the
Lyman limit, carry no information and NaN for masked or empty bands becomes that clip). The PCA fit (a covariance
and its eigenvectors) and the projections run through torch on the GPU: set-up of the training targets, not the
emulator's inference, which is the HIP network alone.
"""
from __future__ import annotations

import math
import time

import numpy as np

from . import hip as H
from . import synth
from .emulator import MLP
from .params import NPAR, SLOT

MAG_CLIP = 30.0
SPEC_FLOOR = 1e-20
# parameters that act logarithmically: their features are log10 of them
LOG_INPUTS = ("tage", "tau", "duste_umin")


class InputSpace:
    """The emulator's inputs: parameter names (slots of opensps_params.h), their box (uniform sampling and the
    serving path's range), whether the batch runs on the SSP operator (per-model IMF slopes), and the features the
    network takes: the inputs (log10 of LOG_INPUTS) and log10(1 + z) after z (the distance acts logarithmically)."""

    def __init__(self, names, box, use_ops=False):
        self.names = list(names)
        self.box = dict(box)
        self.use_ops = bool(use_ops)
        self.iz = self.names.index("zred")
        self.n_feat = len(self.names) + 1

    def sample(self, n, seed):
        """(n, k) inputs uniform in the box."""
        rng = np.random.Generator(np.random.PCG64(seed))
        return np.stack([rng.uniform(*self.box[k], n) for k in self.names], axis=1)

    def bounds(self):
        return (np.array([self.box[k][0] for k in self.names]), np.array([self.box[k][1] for k in self.names]))

    def features(self, x):
        x = np.asarray(x, dtype=np.float64)
        cols = []
        for i, k in enumerate(self.names):
            cols.append(np.log10(x[:, i]) if k in LOG_INPUTS else x[:, i])
            if i == self.iz:
                cols.append(np.log10(1.0 + x[:, i]))
        return np.stack(cols, axis=1).astype(np.float32)

    def to_prm(self, x, base=None):
        """(n, k) inputs -> (NPAR, n) parameter slots (delayed-tau SFH; the others at the context's defaults)."""
        n = x.shape[0]
        p = synth.default_params(n) if base is None else np.repeat(base[:, None], n, axis=1).copy()
        p[SLOT["sfh"]] = 4.0
        for i, k in enumerate(self.names):
            p[SLOT[k]] = x[:, i]
        return p

    def to_prm_device(self, x):
        """The same on the device: x (n, k) tensor -> (NPAR, n) float64 tensor."""
        import torch
        n = x.shape[0]
        base = torch.as_tensor(np.asarray(synth.default_params(1))[:, 0], device=x.device)
        p = base[:, None].repeat(1, n).contiguous()
        p[SLOT["sfh"]] = 4.0
        for i, k in enumerate(self.names):
            p[SLOT[k]] = x[:, i].double()
        return p.contiguous()


BASE = InputSpace(["tage", "tau", "logzsol", "dust2", "dust_index", "zred", "sigma_smooth"],
                  {"tage": (0.1, 13.7), "tau": (0.1, 10.0), "logzsol": (-2.0, 0.3), "dust2": (0.0, 2.0),
                   "dust_index": (-1.2, -0.4), "zred": (0.0, 4.0), "sigma_smooth": (50.0, 300.0)})
# the full input set: BASE, the Kroupa slopes in the operator workload's range and the dust emission / AGN parameters
# over the Draine & Li tables' q_PAH and U_min, a fraction gamma of dust in the power-law U distribution up to 0.5, and
# an AGN of up to half the stars' bolometric luminosity
FULL = InputSpace(BASE.names + ["imf1", "imf2", "imf3", "duste_umin", "duste_qpah", "duste_gamma", "fagn"],
                  dict(BASE.box, imf1=(0.3, 3.3), imf2=(0.3, 3.3), imf3=(0.3, 3.3), duste_umin=(0.1, 25.0),
                       duste_qpah=(0.47, 4.58), duste_gamma=(0.0, 0.5), fagn=(0.0, 0.5)), use_ops=True)

# the base space's forms, as module-level names
INPUTS = BASE.names
BOX = BASE.box
N_FEAT = BASE.n_feat


def features(x):
    """The network's inputs from BASE's parameters (n, 7): log10 tage, log10 tau, logzsol, dust2, dust_index, z,
    log10(1 + z), sigma_smooth."""
    return BASE.features(x)


def distance_term(z, cosmo):
    """The magnitudes' distance term (distance_term, src/spectra/opensps_batch_photometry.f90:784-794): 5 log10(d_L /
    10 pc) - 2.5 log10(1 + z) with d_L linear in the cosmology table; 0 at z = 0. The emulator learns the magnitudes
    without it (a smooth function of the parameters) and adds it back."""
    cz, cdl = (np.asarray(a, dtype=np.float64) for a in cosmo)
    if cz[0] > cz[-1]:
        cz, cdl = cz[::-1], cdl[::-1]
    z = np.asarray(z, dtype=np.float64)
    dm = np.interp(z, cz, cdl)
    out = np.zeros_like(z)
    ok = (np.abs(z) > np.finfo(np.float64).tiny) & (dm > np.finfo(np.float64).tiny)
    out[ok] = 5.0 * np.log10(dm[ok] / 10.0) - 2.5 * np.log10(1.0 + z[ok])
    return out


def sample_inputs(n, seed, space=BASE):
    """(n, k) inputs uniform in the space's box."""
    return space.sample(n, seed)


def to_prm(x, base=None, space=BASE):
    return space.to_prm(x, base)


def to_prm_device(x, space=BASE):
    return space.to_prm_device(x)


def log_spectra(spec):
    """log10 of the spectra with the floor at 1e-20 of each model's peak (torch, float32)."""
    import torch
    peak = spec.amax(dim=1, keepdim=True).clamp_min(1e-38)
    return torch.log10(torch.maximum(spec, peak * SPEC_FLOOR))


def clip_mags(mags):
    """NaN and bands fainter than the model's brightest + MAG_CLIP -> that clip (torch)."""
    import torch
    fin = torch.isfinite(mags)
    bright = torch.where(fin, mags, torch.full_like(mags, float("inf"))).amin(dim=1, keepdim=True)
    lim = bright + MAG_CLIP
    return torch.where(fin, torch.minimum(mags, lim), lim)


def generate(dctx, x, precision=H.MIXED, batch=16384, pca=None, cosmo=None, space=BASE):
    """Runs the batch on the inputs x (n, k) (host) in chunks. Returns device tensors: the log-spectra (n, nspec)
    when pca is None, else their PCA coefficients (n, k), and the clipped magnitudes (n, nbands) less the distance
    term when cosmo is given, float32."""
    import torch
    n = x.shape[0]
    outs_s, outs_m = [], []
    for i0 in range(0, n, batch):
        xb = torch.as_tensor(x[i0:i0 + batch], device="cuda")
        prm = space.to_prm_device(xb)
        o = dctx.batch(prm, precision=precision, want=("mags", "spec"), use_ops=int(space.use_ops))
        ls = log_spectra(o["spec"].float())
        outs_s.append(ls if pca is None else pca.project(ls))
        m = clip_mags(o["mags"])
        if cosmo is not None:
            m = m - torch.as_tensor(distance_term(x[i0:i0 + batch, space.iz], cosmo), device="cuda")[:, None]
        outs_m.append(m.float())
    return torch.cat(outs_s), torch.cat(outs_m)


class PCA:
    """The mean and the leading k eigenvectors of the covariance of log-spectra (torch, on the GPU)."""

    def __init__(self, ls, k=128):
        import torch
        x = ls.double()
        self.mean = x.mean(dim=0)
        xc = x - self.mean
        cov = xc.T @ xc / (x.shape[0] - 1)
        evals, evecs = torch.linalg.eigh(cov)
        order = torch.argsort(evals, descending=True)[:k]
        self.basis = evecs[:, order].float().contiguous()          # (nspec, k)
        self.mean = self.mean.float()
        self.explained = float(evals[order].sum() / evals.sum())

    def project(self, ls):
        return (ls - self.mean) @ self.basis

    def reconstruct(self, c):
        return c @ self.basis.T + self.mean


class Emulator:
    """A trained emulator with the batch as fallback: inputs outside the training box, and batches whose
    spot-check (a random subset recomputed by the batch) has a median magnitude error above tol, go to the batch."""

    def __init__(self, dctx, mlp, pca, n_pca, nbands, box, cosmo, precision=H.MIXED, space=BASE):
        self.dctx, self.mlp, self.pca, self.n_pca, self.nbands, self.box = dctx, mlp, pca, n_pca, nbands, box
        self.cosmo = cosmo
        self.precision = precision
        self.space = space
        self.stats = {"models": 0, "fallback_models": 0, "spot_checked": 0, "batches_failed": 0}

    @staticmethod
    def train(dctx, cosmo, n_train=4_000_000, n_pca_fit=200_000, n_pca=128, epochs=100, batch=8192, lr=2e-3,
              seed=101, hidden=512, layers=4, log=None, precision=H.MIXED, pca_weight=0.1, space=BASE):
        """Trains on n_train samples of the space (the context set up for it: the operator and the IMF when the space
        takes IMF slopes, the emission templates and flags for the dust emission inputs)."""
        import torch
        t0 = time.perf_counter()
        nb = dctx.nbands
        x_fit = space.sample(n_pca_fit, seed)
        ls_fit, _ = generate(dctx, x_fit, precision, space=space)
        pca = PCA(ls_fit, n_pca)
        del ls_fit
        t_pca = time.perf_counter() - t0
        x_tr = space.sample(n_train, seed + 1)
        c_tr, m_tr = generate(dctx, x_tr, precision, pca=pca, cosmo=cosmo, space=space)
        t_data = time.perf_counter() - t0 - t_pca
        xd = torch.as_tensor(space.features(x_tr), device="cuda")
        td = torch.cat([c_tr, m_tr], dim=1).contiguous()
        del c_tr, m_tr
        n_out = n_pca + nb
        mlp = MLP(space.n_feat, n_out, hidden, layers, seed=seed)
        in_mu, in_sd = xd.mean(dim=0), xd.std(dim=0)
        out_mu, out_sd = td.mean(dim=0), td.std(dim=0).clamp_min(1e-6)
        mlp.set_scaling(in_mu.cpu().numpy(), in_sd.cpu().numpy(), out_mu.cpu().numpy(), out_sd.cpu().numpy())
        # the magnitudes are the emulator's product; the PCA coefficients (the spectra) weigh less
        mlp.set_loss_weights(np.concatenate([np.full(n_pca, pca_weight), np.ones(nb)]).astype(np.float32))
        steps_per_epoch = n_train // batch
        total = epochs * steps_per_epoch
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed)
        step = 0
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n_train, device="cuda", generator=gen)
            for s in range(steps_per_epoch):
                idx = perm[s * batch:(s + 1) * batch]
                xb = xd[idx].contiguous()
                tb = td[idx].contiguous()
                lr_t = 0.5 * lr * (1.0 + math.cos(math.pi * step / total))   # cosine decay
                want = s == steps_per_epoch - 1
                l = mlp.train_step(xb, tb, lr_t, want_loss=want)
                step += 1
                if want:
                    hist.append(l)
                    if log:
                        log(f"epoch {ep + 1}/{epochs}: loss {l:.5f}")
        mlp.sync()
        t_train = time.perf_counter() - t1
        emu = Emulator(dctx, mlp, pca, n_pca, nb, space.bounds(), cosmo, precision, space)   # the sampled box
        emu.train_info = {"inputs": space.names, "n_feat": space.n_feat, "n_train": n_train, "n_pca_fit": n_pca_fit, "pca_explained": pca.explained, "epochs": epochs,
                          "batch": batch, "lr": lr, "steps": total, "loss_history": hist, "pca_s": t_pca,
                          "data_s": t_data, "train_s": t_train,
                          "train_samples_per_s": total * batch / t_train}
        return emu

    def predict_raw(self, f_dev, dist_dev=None):
        """The network alone: device features (n, N_FEAT) float32 -> (PCA coefficients (n, k), magnitudes
        (n, nbands); the distance term dist_dev (n) added when given)."""
        y = self.mlp.predict(f_dev)
        m = y[:, self.n_pca:]
        if dist_dev is not None:
            m = m + dist_dev[:, None]
        return y[:, :self.n_pca], m

    def inputs(self, x):
        """Host parameters (n, k) -> device features and distance terms."""
        import torch
        return (torch.as_tensor(self.space.features(x), device="cuda"),
                torch.as_tensor(distance_term(x[:, self.space.iz], self.cosmo).astype(np.float32), device="cuda"))

    def in_box(self, x):
        lo, hi = self.box
        return np.all((x >= lo) & (x <= hi), axis=1)

    def predict(self, x, spot=0.01, tol=0.01, seed=0):
        """Magnitudes (n, nbands) of the inputs x (n, k) (host): the emulator inside the box, the batch outside it
        and for the whole batch when the spot-check's median |dmag| exceeds tol. Returns (mags, info)."""
        import torch
        n = x.shape[0]
        inb = self.in_box(x)
        _, mags = self.predict_raw(*self.inputs(x))
        mags = mags.double()
        rng = np.random.default_rng(seed)
        idx_in = np.nonzero(inb)[0]
        k = min(len(idx_in), max(16, int(spot * n)))
        failed = False
        med = 0.0
        if k > 0:
            pick = rng.choice(idx_in, k, replace=False)
            ref = self._batch(x[pick])
            d = (mags[torch.as_tensor(pick, device="cuda")] - ref).abs()
            med = float(d[torch.isfinite(d)].median()) if torch.isfinite(d).any() else 0.0
            failed = med > tol
        fb = np.ones(n, bool) if failed else ~inb
        if fb.any():
            mags[torch.as_tensor(np.nonzero(fb)[0], device="cuda")] = self._batch(x[fb])
        self.stats["models"] += n
        self.stats["fallback_models"] += int(fb.sum())
        self.stats["spot_checked"] += k
        self.stats["batches_failed"] += int(failed)
        return mags, {"spot_median_dmag": med, "fallback": int(fb.sum()), "spot_failed": failed}

    def _batch(self, x):
        import torch
        prm = self.space.to_prm_device(torch.as_tensor(x, device="cuda"))
        o = self.dctx.batch(prm, precision=self.precision, want=("mags",), use_ops=int(self.space.use_ops))
        return clip_mags(o["mags"])

    def fallback_rate(self):
        return self.stats["fallback_models"] / max(1, self.stats["models"])


def validate(emu, dctx, n=16384, seed=999, mag_range=20.0):
    """The batch as the judge: |dmag| of the emulator against the FP64 batch on a validation set, over the bands
    within mag_range of each model's brightest. Returns the median, 95th percentile and max, and the spectra's
    reconstruction error (rms dex of the log-spectrum over the pixels within 10 dex of the peak)."""
    import torch
    sp = emu.space
    x = sp.sample(n, seed)
    c, m = emu.predict_raw(*emu.inputs(x))
    o = dctx.batch(sp.to_prm_device(torch.as_tensor(x, device="cuda")), precision=H.FP64, want=("mags", "spec"),
                   use_ops=int(sp.use_ops))
    ref = o["mags"]
    fin = torch.isfinite(ref)
    bright = torch.where(fin, ref, torch.full_like(ref, float("inf"))).amin(dim=1, keepdim=True)
    ok = fin & (ref - bright <= mag_range)
    d = (m.double() - ref).abs()[ok]
    ls = log_spectra(o["spec"].float())
    rec = emu.pca.reconstruct(c)
    near = ls >= ls.amax(dim=1, keepdim=True) - 10.0
    rms = float(torch.sqrt(((rec - ls)[near] ** 2).mean()))
    pca_only = float(torch.sqrt(((emu.pca.reconstruct(emu.pca.project(ls)) - ls)[near] ** 2).mean()))
    return {"n": n, "median_dmag": float(d.median()), "p95_dmag": float(torch.quantile(d.float()[:1 << 24], 0.95)),
            "max_dmag": float(d.max()), "bands_within": mag_range, "spec_rms_dex": rms, "pca_floor_rms_dex": pca_only}
