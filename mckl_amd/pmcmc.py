"""Particle marginal Metropolis-Hastings on the device log-likelihood.

``PMCMCMutation`` follows PMCMCMutation<Param, T> (algorithm/pmcmc.hpp:
71-158) step for step: at iteration 0 it runs the particle filter at the
current parameter and keeps its log normalising constant; afterwards it
starts from -prior(param) - log_nc, lets each update object move a copy p of
the parameter (adding its log proposal ratio), adds prior(p), runs the
filter at p, adds its log normalising constant, and accepts when
log(U01(rng)) is below the total (U01Distribution<double> is
U01CanonicalDistribution<double> in the reference's default configuration,
random/u01_distribution.hpp:42,225-233).  The filter is a one-filter
SMCBatch (a single persistent launch per run) whose log-likelihood is the
reference's log_nc.

``GenericPMCMCMutation`` is PMCMCMutation over a GenericSMCSampler with user
evaluation objects (any model), mirroring the reference's loop exactly.

``PMCMCBatch`` advances C independent chains of the same kind together:
one batched SMC launch evaluates all C proposals, the acceptance uniforms
of all chains come from one device fill.  It has no counterpart in the
reference, which runs one chain at a time.

Parameters are float64 arrays of the built-in models' theta (LG: a, sx, sy;
SV: phi, sigma, beta).  A proposal whose prior is -inf is rejected without
running the filter (the built-in models reject such parameters).  Filter
run r (counted from 0 over the object's life) of chain c uses the SMC seed
seed + 2^32 * (r * C + c + 1), so no two runs share particle streams.
"""
import math

import numpy as np

from . import (SMCBatch, Philox4x32, U01CanonicalDistribution,
               NormalDistribution, STRATIFIED, MCKLError, rand)


def _u01_canonical(rng, n):
    import torch
    out = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
    rand(rng, U01CanonicalDistribution("float64"), n, out)
    return out[:n].cpu().numpy()


def _normals(rng, n):
    import torch
    out = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
    rand(rng, NormalDistribution(0.0, 1.0, "float64"), n, out)
    return out[:n].cpu().numpy()


def random_walk(scales):
    """Update object: Gaussian random walk p += scales * z, z ~ N(0, I)
    drawn with NormalDistribution<double> from the chain's engine; log
    proposal ratio 0.  Works on one parameter vector (PMCMCMutation) and on
    a (C, 3) array (PMCMCBatch, one fill for all chains)."""
    s = np.asarray(scales, np.float64)

    def update(rng, p):
        z = _normals(rng, p.size).reshape(p.shape)
        p += s * z
        return np.zeros(p.shape[:-1]) if p.ndim > 1 else 0.0

    return update


def lognormal_walk(index, sd):
    """Update object of the reference's PMCMC example
    (example/algorithm/include/algorithm_pmcmc.hpp:121-134): multiply
    component `index` by u = exp(sd * z) and return log(u)."""

    def update(rng, p):
        z = _normals(rng, p.size // 3 if p.ndim > 1 else 1)
        logu = sd * z
        if p.ndim > 1:
            p[:, index] *= np.exp(logu)
            return logu
        p[index] *= math.exp(float(logu[0]))
        return float(logu[0])

    return update


class PMCMCMutation:
    """PMCMCMutation (algorithm/pmcmc.hpp:71-158) for the built-in models.

    N particles, M filter iterations (the first M observations of y), prior
    a callable param -> log density.  ``update(eval)`` adds an update object
    eval(rng, p) -> log proposal ratio that moves p in place.
    ``mutation(iter, param)`` moves param in place and returns 1 if the
    proposal was accepted."""

    def __init__(self, N, M, prior, model, y, seed=1, resample_threshold=0.5,
                 scheme=STRATIFIED):
        y = np.ascontiguousarray(y, np.float64)[:M]
        if len(y) != M:
            raise MCKLError("PMCMCMutation: fewer than M observations")
        self.M = M
        self.N = N
        self._prior = prior
        self._eval = []
        self._seed = int(seed) & 0xFFFFFFFF
        self._runs = 0
        self.rng = Philox4x32(seed)
        init = (0.5, 1.0, 1.0)
        self.pf = SMCBatch(model, [N], [init], y[None, :], [self._seed],
                           resample_threshold, scheme)
        self.log_nc = 0.0

    def prior(self, prior):
        self._prior = prior

    def update(self, ev):
        self._eval.append(ev)
        return len(self._eval) - 1

    def reset(self):
        self._eval = []

    def _loglik(self, p):
        seed = self._seed + ((self._runs + 1) << 32)
        self._runs += 1
        self.pf.set_params(np.asarray(p, np.float64)[None, :], [seed])
        self.pf.run()
        return float(self.pf.log_likelihood()[0])

    def __call__(self, iter, param):
        if iter == 0:
            self.log_nc = self._loglik(param)
            return 0
        lnc = self.log_nc
        prob = -self._prior(param) - lnc
        p = np.array(param, np.float64)
        for ev in self._eval:
            prob += ev(self.rng, p)
        prob += self._prior(p)
        new = -math.inf
        if prob > -math.inf:
            new = self._loglik(p)
            prob += new
        else:
            self._runs += 1  # the run's seed stays reserved
        u = math.log(float(_u01_canonical(self.rng, 1)[0]))
        if u < prob:
            param[...] = p
            self.log_nc = new
            return 1
        return 0


class PMCMCBatch:
    """C PMMH chains in lockstep.  prior maps a (C, 3) array to C log
    densities; update objects eval(rng, P) move the (C, 3) array P in place
    and return C log proposal ratios.  ``mutation(iter, params)`` moves the
    (C, 3) array params in place and returns the C acceptance flags."""

    def __init__(self, N, M, prior, model, y, nchains, seed=1,
                 resample_threshold=0.5, scheme=STRATIFIED):
        y = np.ascontiguousarray(y, np.float64)[:M]
        if len(y) != M:
            raise MCKLError("PMCMCBatch: fewer than M observations")
        self.C = int(nchains)
        self.M = M
        self._prior = prior
        self._eval = []
        self._seed = int(seed) & 0xFFFFFFFF
        self._runs = 0
        self.rng = Philox4x32(seed)
        init = np.tile([0.5, 1.0, 1.0], (self.C, 1))
        seeds = np.full(self.C, self._seed, np.uint64)
        self.pf = SMCBatch(model, np.full(self.C, N, np.uint32), init,
                           np.tile(y, (self.C, 1)), seeds, resample_threshold,
                           scheme)
        self.log_nc = np.zeros(self.C)

    def update(self, ev):
        self._eval.append(ev)
        return len(self._eval) - 1

    def _loglik(self, P, ok):
        # rows that are not run keep a valid placeholder parameter
        Q = np.where(ok[:, None], P, np.array([0.5, 1.0, 1.0]))
        base = self._runs * self.C
        self._runs += 1
        seeds = (self._seed + ((base + 1 + np.arange(self.C, dtype=np.uint64))
                               << np.uint64(32))).astype(np.uint64)
        self.pf.set_params(Q, seeds)
        self.pf.run()
        ll = self.pf.log_likelihood()
        return np.where(ok, ll, -np.inf)

    def __call__(self, iter, params):
        C = self.C
        if iter == 0:
            self.log_nc = self._loglik(params, np.ones(C, bool))
            return np.zeros(C, np.int32)
        lnc = self.log_nc
        prob = -np.asarray(self._prior(params), np.float64) - lnc
        P = np.array(params, np.float64)
        for ev in self._eval:
            prob = prob + np.asarray(ev(self.rng, P), np.float64)
        prob = prob + np.asarray(self._prior(P), np.float64)
        ok = prob > -np.inf
        new = self._loglik(P, ok)
        prob = prob + new
        u = np.log(_u01_canonical(self.rng, C))
        acc = u < prob
        params[acc] = P[acc]
        self.log_nc = np.where(acc, new, lnc)
        return acc.astype(np.int32)


class GenericPMCMCMutation:
    """PMCMCMutation<Param, T> (algorithm/pmcmc.hpp:71-158) over a
    GenericSMCSampler with user evaluation objects, step for step.

    ``pf()`` is the particle filter: its particle is a PMCMCParticle, whose
    ``param`` the evaluation objects read and whose log normalising
    constant they accumulate with ``add_log_nc``.  M: filter iterations per
    run.  ``update(eval)`` adds eval(rng, p) -> log proposal ratio, moving
    the parameter array p in place with the particle's engine.  At
    iteration 0 the filter runs at the starting parameter; afterwards the
    proposal's filter always runs (the prior is added, not tested, as in
    the reference), the acceptance draw is log(U01Canonical<double>) from
    the particle's engine, and a rejected proposal restores the parameter
    and log normalising constant but keeps the particles, as the reference
    does.  ``mutation(iter, param)`` moves param in place and returns 1 if
    the proposal was accepted."""

    def __init__(self, N, M, prior, dim=1, dtype=None, rng=None, seed=1):
        from .sampler import GenericSMCSampler, PMCMCParticle
        self.M = int(M)
        self._prior = prior
        self._eval = []
        part = PMCMCParticle(N, dim, dtype, rng, seed)
        self._pf = GenericSMCSampler(N, particle=part)

    def pf(self):
        return self._pf

    def prior(self, prior):
        self._prior = prior

    def update(self, ev):
        self._eval.append(ev)
        return len(self._eval) - 1

    def reset(self):
        self._eval = []
        self._pf.reset()

    def _run(self, p):
        part = self._pf.particle()
        self._pf.clear()
        part.reset(p, 0.0)
        self._pf.iterate(self.M)

    def __call__(self, iter, param):
        part = self._pf.particle()
        if iter == 0:
            self._run(param)
            return 0
        lnc = part.log_nc()
        prob = -self._prior(param) - lnc
        p = np.array(param, np.float64)
        for ev in self._eval:
            prob += ev(part.rng(), p)
        prob += self._prior(p)
        self._run(p)
        prob += part.log_nc()
        u = math.log(float(_u01_canonical(part.rng(), 1)[0]))
        if u < prob:
            param[...] = p
            return 1
        part.reset(param, lnc)
        return 0
