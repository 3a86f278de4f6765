"""SMCSampler<T> with evaluation objects on device particles.

The reference's generic sampler (algorithm/smc.hpp:151-421) runs user
evaluation objects eval(iter, particle) in three slots -- selection,
resample, mutation -- and estimators after each slot.  Here the particle
system lives in device memory and the library does what the reference's
sampler does itself: the weights (Weight on device, FP64), the ESS test,
the built-in resampling schemes (ResampleEval over the device resampling
kernels, drawing from Particle::rng()) and the selection of the states.
User evaluation objects are host callables that enqueue their own device
work (torch operations or their own kernels) on the current stream.

``Particle``        Particle<T> (core/particle.hpp:277-360): state (N x dim
                    torch tensor, StateMatrix<RowMajor, Dynamic, T>, float32
                    or float64), weight(), rng(), select()
``ResampleEval``    ResampleEval<T> (algorithm/resample.hpp:470-511)
``SMCEstimator``    SMCEstimator<T, double> (algorithm/smc.hpp:48-136); the
                    weighted sums are the device's exactly rounded sums
                    (the reference calls an external BLAS dgemv)
``GenericSMCSampler``  SMCSampler<T> (algorithm/smc.hpp:151-421)
``PMCMCParticle``   a Particle carrying PMCMCStateMatrix's parameter and log
                    normalising constant (algorithm/pmcmc.hpp:46-66)

The ESS test reads the ESS to the host once per iteration, as the
reference does; the built-in-model samplers (SMCSampler, SMCBatch) take
the decision on the device instead.
"""
import math

import numpy as np

from . import (Weight, Philox4x32, MCKLError, check, lib, _ptr, _stream,
               ResampleMultinomial, ResampleStratified, ResampleSystematic,
               ResampleResidual, ResampleResidualStratified,
               ResampleResidualSystematic, RowMajor, ColMajor, estimate)

_SCHEMES = {0: ResampleMultinomial, 1: ResampleStratified,
            2: ResampleSystematic, 3: ResampleResidual,
            4: ResampleResidualStratified, 5: ResampleResidualSystematic}


def _torch():
    import torch
    return torch


class Particle:
    """Particle<T> (core/particle.hpp:277-360) in device memory.

    ``state`` is an (N, dim) tensor on the GPU (row-major: particle i's
    values are row i), ``weight()`` the device Weight, ``rng()`` the engine
    ResampleEval draws from (Philox4x32(seed) unless one is given)."""

    def __init__(self, N, dim=1, dtype=None, rng=None, seed=1):
        torch = _torch()
        dtype = torch.float64 if dtype is None else dtype
        if dtype not in (torch.float32, torch.float64):
            raise MCKLError("Particle: state type must be float32 or float64")
        if N < 1 or dim < 1:
            raise MCKLError("Particle: N and dim must be positive")
        self.state = torch.zeros((N, dim), dtype=dtype, device="cuda")
        self._spare = torch.empty_like(self.state)
        self._weight = Weight(N)
        self._rng = rng if rng is not None else Philox4x32(seed)

    def size(self):
        return self.state.shape[0]

    def dim(self):
        return self.state.shape[1]

    def weight(self):
        return self._weight

    def rng(self):
        return self._rng

    def select(self, n, idx, validate=True):
        """Particle::select (core/particle.hpp:324-331): row j of the new
        state is row idx[j] of the old one (j < n; idx an int32 device
        tensor), then equal weights of size n.  With validate (the default)
        the row numbers are verified on the host first, so that a bad index
        raises instead of reading outside the state; ResampleEval, whose
        indices come from the library, skips it."""
        torch = _torch()
        if idx.numel() < n or idx.dtype != torch.int32 or not idx.is_cuda:
            raise MCKLError("Particle.select: idx must be an int32 device "
                            "tensor of at least n entries")
        if validate and n > 0:
            lo, hi = torch.aminmax(idx[:n])
            if int(lo) < 0 or int(hi) >= self.size():
                raise MCKLError("Particle.select: row numbers must lie in "
                                "[0, %d)" % self.size())
        resize = n != self.size()
        dst = (torch.empty((n, self.dim()), dtype=self.state.dtype,
                           device="cuda") if resize else self._spare)
        fn = (lib.mckl_hip_select_f32 if self.state.dtype == torch.float32
              else lib.mckl_hip_select_f64)
        check(fn(n, self.dim(), _ptr(idx), _ptr(self.state), _ptr(dst),
                 _stream()))
        if resize:
            self.state = dst
            self._spare = torch.empty_like(dst)
            self._weight = Weight(n)
        else:
            self._spare, self.state = self.state, dst
            self._weight.set_equal()


class PMCMCParticle(Particle):
    """A Particle whose state also carries PMCMCStateMatrix's parameter
    and log normalising constant (algorithm/pmcmc.hpp:46-66): ``param``,
    ``log_nc()``, ``add_log_nc(nc)`` (a float, or a 0-d device tensor that
    is summed on the device and read by log_nc()), ``reset(p, nc)``."""

    def __init__(self, N, dim=1, dtype=None, rng=None, seed=1):
        super().__init__(N, dim, dtype, rng, seed)
        self.param = None
        self._nc = 0.0
        self._dev = None

    def log_nc(self):
        return self._nc + (float(self._dev.item()) if self._dev is not None
                           else 0.0)

    def add_log_nc(self, nc):
        if hasattr(nc, "is_cuda"):
            self._dev = nc if self._dev is None else self._dev + nc
        else:
            self._nc += float(nc)

    def reset(self, p, nc=0.0):
        self.param = np.array(p, np.float64)
        self._nc = float(nc)
        self._dev = None


class ResampleEval:
    """ResampleEval<T> (algorithm/resample.hpp:470-511): replication
    numbers of a resampling algorithm from Particle::rng() and the weights,
    the stable rep -> index transform, Particle::select.  ``alg`` is a
    Resample* object or a ResampleScheme number."""

    def __init__(self, alg):
        if isinstance(alg, (int, np.integer)):
            if int(alg) not in _SCHEMES:
                raise MCKLError("**SMCSampler::resample** used with unknown "
                                "resample scheme")
            alg = _SCHEMES[int(alg)]()
        self._alg = alg
        self._rep = self._idx = None

    def __call__(self, it, particle):
        torch = _torch()
        N = particle.size()
        if self._idx is None or self._idx.numel() != N:
            self._rep = torch.empty(N, dtype=torch.int32, device="cuda")
            self._idx = torch.empty(N, dtype=torch.int32, device="cuda")
        self._alg(N, N, particle.rng(), particle.weight().data, self._rep,
                  self._idx)
        particle.select(N, self._idx, validate=False)


class SMCEstimator:
    """SMCEstimator<T, double> (algorithm/smc.hpp:48-136).

    ``eval(iter, dim, particle, u)`` fills the float64 device tensor u:
    (N, dim) for RowMajor, (dim, N) for ColMajor, or the (dim,) estimate
    itself when record_only.  Otherwise estimate() appends
    r_j = sum_i W_i u(i, j) (each sum exactly rounded) as the next row."""

    def __init__(self, dim, eval=None, layout=RowMajor, record_only=False):
        if dim < 1:
            raise MCKLError("SMCEstimator: dim must be positive")
        self._dim = int(dim)
        self._eval = eval
        self._layout = layout
        self._record_only = bool(record_only)
        self._rows = []
        self._u = None

    def dim(self):
        return self._dim

    def num_iter(self):
        return len(self._rows)

    def record_only(self):
        return self._record_only

    def set_eval(self, eval, layout=RowMajor, record_only=False):
        """SMCEstimator::estimate(eval, layout, record_only) (smc.hpp:76-84)."""
        self._eval = eval
        self._layout = layout
        self._record_only = bool(record_only)

    def clear(self):
        self._rows = []

    def reserve(self, n):
        pass

    def estimate(self, it, particle):
        torch = _torch()
        if self._eval is None:
            raise MCKLError("**Estimator::eval** used with an invalid "
                            "evaluation object")
        d = self._dim
        if self._record_only:
            r = torch.empty(d, dtype=torch.float64, device="cuda")
            self._eval(it, d, particle, r)
            self._rows.append(r)
            return
        n = particle.size()
        shape = (n, d) if self._layout == RowMajor else (d, n)
        if self._u is None or tuple(self._u.shape) != shape:
            self._u = torch.empty(shape, dtype=torch.float64, device="cuda")
        self._eval(it, d, particle, self._u)
        self._rows.append(estimate(self._u.reshape(-1), particle.weight().data,
                                   d, self._layout))

    def estimates(self):
        """(num_iter, dim) array of the stored estimates."""
        torch = _torch()
        if not self._rows:
            return np.zeros((0, self._dim))
        return torch.stack(self._rows).cpu().numpy()

    def read_estimate(self, j):
        """Variable j over the iterations (EstimateMatrix::variable)."""
        return self.estimates()[:, j].copy()

    def average(self, cut=0, thin=1):
        """Estimator::average (core/estimator.hpp:62-93): the rows cut,
        cut + thin, ... summed and, as the reference does, divided by the
        number of stored iterations; when fewer than thin rows follow cut,
        row cut itself."""
        e = self.estimates()
        n = e.shape[0]
        if cut >= n:
            return np.zeros(0)
        thin = max(int(thin), 1)
        if n - cut < thin:
            return e[cut].copy()
        s = np.zeros(self._dim)
        for i in range(cut, n, thin):
            s = s + e[i]
        return s * (1.0 / n)


class GenericSMCSampler:
    """SMCSampler<T> (algorithm/smc.hpp:151-421) on a device Particle.

    Evaluation objects are callables eval(iter, particle).  iterate(n) runs
    do_iterate (smc.hpp:389-406) n times: selection evals, selection
    estimators, the size and ESS recorded, resample evals when
    ESS < N * threshold, resample estimators, mutation evals, mutation
    estimators.  The threshold starts at resample_threshold_never()."""

    def __init__(self, N, dim=1, dtype=None, rng=None, seed=1, particle=None):
        self._particle = (particle if particle is not None
                          else Particle(N, dim, dtype, rng, seed))
        self._evals = [[], [], []]
        self._ests = [[], [], []]
        self._iter = 0
        self._threshold = -math.inf
        self._size_history = []
        self._ess_history = []

    @staticmethod
    def resample_threshold_never():
        return -math.inf

    @staticmethod
    def resample_threshold_always():
        return math.inf

    def size(self):
        return self._particle.size()

    def particle(self):
        return self._particle

    def num_iter(self):
        return len(self._size_history)

    def resample_threshold(self, threshold=None):
        if threshold is None:
            return self._threshold
        self._threshold = float(threshold)

    def _add(self, slot, ev):
        if not callable(ev):
            raise MCKLError("SMCSampler: evaluation objects are callables")
        self._evals[slot].append(ev)
        return len(self._evals[slot]) - 1

    def selection(self, ev):
        return self._add(0, ev)

    def resample(self, ev):
        """A callable, or a ResampleScheme number for the built-in
        ResampleEval (smc.hpp:246-270)."""
        if isinstance(ev, (int, np.integer)):
            ev = ResampleEval(int(ev))
        return self._add(1, ev)

    def mutation(self, ev):
        return self._add(2, ev)

    def _add_est(self, slot, est):
        if not isinstance(est, SMCEstimator):
            raise MCKLError("SMCSampler: estimators are SMCEstimator objects")
        self._ests[slot].append(est)
        return len(self._ests[slot]) - 1

    def selection_estimator(self, est):
        if isinstance(est, (int, np.integer)):
            return self._ests[0][est]
        return self._add_est(0, est)

    def resample_estimator(self, est):
        if isinstance(est, (int, np.integer)):
            return self._ests[1][est]
        return self._add_est(1, est)

    def mutation_estimator(self, est):
        if isinstance(est, (int, np.integer)):
            return self._ests[2][est]
        return self._add_est(2, est)

    def reserve(self, n):
        pass

    def clear(self):
        """Clear all history (smc.hpp:205-212)."""
        for slot in self._ests:
            for e in slot:
                e.clear()
        self._iter = 0
        self._size_history = []
        self._ess_history = []

    def reset(self):
        """Remove every evaluation object and estimator, clear the history
        (smc.hpp:197-202)."""
        self._evals = [[], [], []]
        self._ests = [[], [], []]
        self.clear()

    def iterate(self, n=1):
        for _ in range(int(n)):
            self._do_iterate()

    def _do_iterate(self):
        p = self._particle
        for ev in self._evals[0]:
            ev(self._iter, p)
        self._estimate(0)
        self._size_history.append(p.size())
        ess = p.weight().ess()
        self._ess_history.append(ess)
        if ess < p.size() * self._threshold:
            for ev in self._evals[1]:
                ev(self._iter, p)
        self._estimate(1)
        for ev in self._evals[2]:
            ev(self._iter, p)
        self._estimate(2)
        self._iter += 1

    def _estimate(self, slot):
        for est in self._ests[slot]:
            est.estimate(self._iter, self._particle)

    def size_history(self):
        return np.array(self._size_history, np.uint64)

    def ess_history(self):
        return np.array(self._ess_history, np.float64)

    def summary(self, layout=RowMajor):
        """Sampler::summary (core/sampler.hpp:80-118): every estimator's
        estimates side by side, num_iter rows; ColMajor returns the
        transpose's memory order (column after column)."""
        cols = [e.estimates() for slot in self._ests for e in slot]
        if not cols:
            return np.zeros((self.num_iter(), 0))
        m = np.concatenate(cols, axis=1)
        return np.asfortranarray(m) if layout == ColMajor else m
