"""FastGradientAscent, the reference's default estimator, driving a model whose
log-likelihood and gradient run on the GPU.

The loops restate src/core/estimators/FastGradientAscent.cxx:50-305 (Nesterov-accelerated
gradient ascent with per-block step sizes, an Armijo-like acceptance test and a
multivariate line search) and GradientAscent.cxx:40-238 over the model interface of
AbstractStatisticalModel, making the model calls in the reference's order:

  model.GetFixedEffects() -> dict
  model.SetFixedEffects(fixed)
  model.UpdateFixedEffectsAndComputeCompleteLogLikelihood(pop_rer, ind_rer)
      -> (terms: float64 array of 2, out_of_box: bool)
  model.ComputeCompleteLogLikelihoodGradient(pop_rer, ind_rer)
      -> (pop_grad: dict, ind_grad: dict of lists, grad_sq_norms: float64 array)

The order matters: the reference's atlases move their template geometry one
SetFixedEffects call behind the template fixed effect (AbstractAtlas.cxx:52-63), and the
models here reproduce that.  Every SetFixedEffects of the reference is made, including the
ones around its per-iteration Write (SaveEveryNIters = 1).

Fixed effects and population random effects are dicts name -> array or list of arrays
(std::map, iterated in key order); individual random effects are dicts name -> list of
per-subject arrays.  ``Update`` returns the snapshot the reference saves at the end
(FastGradientAscent.cxx:293-301, GradientAscent.cxx:230-236) as a dict with the field
names of ``io.state.pop_fast_gradient_ascent`` / ``pop_gradient_ascent``.
"""
from __future__ import annotations

import math

import numpy as np


def _copy(effects):
    return {k: ([np.array(x, dtype=np.float64) for x in v] if isinstance(v, list)
                else np.array(v, dtype=np.float64)) for k, v in effects.items()}


def _keys(d):
    return sorted(d, key=lambda k: k.encode())


def _axpy(x, g, c):
    """x + g * c for a matrix or a matrix list (linear-variable arithmetic)."""
    if isinstance(x, list):
        return [a + b * c for a, b in zip(x, g)]
    return x + g * c


def _evaluate(model, fixed, pop, ind):
    """SetFixedEffects then the log-likelihood terms, as every estimator proposal does."""
    model.SetFixedEffects(fixed)
    terms, oob = model.UpdateFixedEffectsAndComputeCompleteLogLikelihood(pop, ind)
    return np.asarray(terms, dtype=np.float64), bool(oob)


def _extrapolate(new, old, c):
    """new + (new - old) * c, elementwise over matrix lists."""
    if isinstance(new, list):
        return [n + (n - o) * c for n, o in zip(new, old)]
    return new + (new - old) * c


class FastGradientAscent:
    def __init__(self, model, *, max_iterations=100, max_line_search_iterations=10,
                 initial_step_size=0.001, step_expand=1.2, step_shrink=0.5,
                 adaptive_tolerance=1e-4, nb_iter_freeze=3, multivariate_line_search=True,
                 callback=None):
        # defaults: SparseDiffeoParameters.cxx:74-89, FastGradientAscent.cxx:22
        self.model = model
        self.max_iterations = int(max_iterations)
        self.max_line_search_iterations = int(max_line_search_iterations)
        self.initial_step_size = float(initial_step_size)
        self.expand = float(step_expand)
        self.shrink = float(step_shrink)
        self.tolerance = float(adaptive_tolerance)
        self.nb_iter_freeze = int(nb_iter_freeze)
        self.multivariate = bool(multivariate_line_search)
        self.callback = callback

    # GradientAscentStep (FastGradientAscent.cxx:313-339, GradientAscent.cxx:242-267)
    @staticmethod
    def _step(fixed, pop, ind, pop_grad, ind_grad, step):
        new_fixed, new_pop, new_ind = _copy(fixed), _copy(pop), _copy(ind)
        k = 0
        for name in _keys(pop_grad):
            if name in new_fixed:
                new_fixed[name] = _axpy(fixed[name], pop_grad[name], step[k])
            else:
                new_pop[name] = _axpy(pop[name], pop_grad[name], step[k])
            k += 1
        for name in _keys(ind_grad):
            new_ind[name] = _axpy(ind[name], ind_grad[name], step[k])
            k += 1
        return new_fixed, new_pop, new_ind

    def _eval(self, fixed, pop, ind):
        return _evaluate(self.model, fixed, pop, ind)

    def Update(self, population_rer=None, individual_rer=None):
        m = self.model
        fixed = _copy(m.GetFixedEffects())
        pop_rer = _copy(population_rer or {})
        ind_rer = _copy(individual_rer or {})
        aux_fixed, aux_pop, aux_ind = _copy(fixed), _copy(pop_rer), _copy(ind_rer)
        history = [np.zeros(0) for _ in range(self.max_iterations + 1)]
        terms, _ = m.UpdateFixedEffectsAndComputeCompleteLogLikelihood(aux_pop, aux_ind)
        history[0] = np.asarray(terms, dtype=np.float64)
        lsq_ref = float(history[0].sum())
        pop_grad, ind_grad, grad_sq = m.ComputeCompleteLogLikelihoodGradient(aux_pop, aux_ind)
        nb_params = len(pop_grad) + len(ind_grad)
        step = np.full(nb_params, self.initial_step_size)
        tau = 1.0
        freeze = 1
        it = 1
        iter_ref = 0
        while it < self.max_iterations + 1:  # FastGradientAscent.cxx:128-283
            found = False
            for li in range(self.max_line_search_iterations):
                new_fixed, new_pop, new_ind = self._step(aux_fixed, aux_pop, aux_ind, pop_grad,
                                                         ind_grad, step)
                new_terms, oob = self._eval(new_fixed, new_pop, new_ind)
                q = float(np.sum(new_terms)) - lsq_ref - float(np.dot(grad_sq, step)) * self.shrink
                if q >= 0 and not oob:
                    found = True
                    if li == 0:
                        freeze += 1
                    break
                if nb_params > 1 and self.multivariate:
                    freeze = 1
                    step = step * self.shrink
                    props = []
                    qs = np.zeros(nb_params)
                    for k in range(nb_params):
                        local = step.copy()
                        local[k] /= self.shrink
                        prop = self._step(aux_fixed, aux_pop, aux_ind, pop_grad, ind_grad, local)
                        t_k, oob_k = self._eval(*prop)
                        props.append((prop, t_k, oob_k))
                        qs[k] = (float(np.sum(t_k)) - lsq_ref
                                 - float(np.dot(grad_sq, local)) * self.shrink)
                    idx = int(np.argmax(qs))  # arma index_max: first maximum
                    if qs[idx] >= 0 and not props[idx][2]:
                        (new_fixed, new_pop, new_ind), new_terms, _ = props[idx]
                        step[idx] /= self.shrink
                        found = True
                        break
                else:
                    freeze = 1
                    step = step * self.shrink
            if not found:
                m.SetFixedEffects(fixed)  # line search exhausted (cxx:207-211)
                break
            history[it] = new_terms
            tau_next = (1.0 + math.sqrt(1.0 + 4.0 * tau * tau)) / 2.0
            if freeze > self.nb_iter_freeze:
                step = step * self.expand
                tau_next = (1.0 + math.sqrt(1.0 + 4.0 * tau * tau / self.expand)) / 2.0
            tau_scale = (tau - 1.0) / tau_next
            for name in _keys(pop_grad):
                if name in aux_fixed:
                    aux_fixed[name] = _extrapolate(new_fixed[name], fixed[name], tau_scale)
                else:
                    aux_pop[name] = _extrapolate(new_pop[name], pop_rer[name], tau_scale)
            for name in _keys(ind_grad):
                aux_ind[name] = _extrapolate(new_ind[name], ind_rer[name], tau_scale)
            fixed, pop_rer, ind_rer = new_fixed, new_pop, new_ind
            tau = tau_next
            terms, _ = self._eval(aux_fixed, aux_pop, aux_ind)
            lsq_ref = float(np.sum(terms))
            m.SetFixedEffects(fixed)      # around the per-iteration Write (cxx:256-260)
            m.SetFixedEffects(aux_fixed)
            pop_grad, ind_grad, grad_sq = m.ComputeCompleteLogLikelihoodGradient(aux_pop, aux_ind)
            if self.callback:
                self.callback(it, history[it])
            d_cur = history[it - 1].sum() - history[it].sum()
            d_ref = history[iter_ref].sum() - history[it].sum()
            if abs(d_cur) < self.tolerance * abs(d_ref):
                break  # tolerance met (cxx:268-271); the loop's increment does not run
            it += 1
        m.SetFixedEffects(fixed)  # cxx:285
        return {
            "computation_end_state": True, "iter": it, "freeze_direction_counter": freeze,
            "tau": tau, "step": step, "grad_sq_norms": np.asarray(grad_sq, dtype=np.float64),
            "log_likelihood_history": history, "fixed_effects": fixed,
            "aux_fixed_effects": aux_fixed, "population_rer": pop_rer,
            "aux_population_rer": aux_pop, "individual_rer": ind_rer,
            "aux_individual_rer": aux_ind, "pop_grad": pop_grad, "ind_grad": ind_grad,
        }


class GradientAscent:
    """src/core/estimators/GradientAscent.cxx:40-238: gradient ascent with per-block steps,
    the same multivariate line search as FastGradientAscent without its Armijo term or
    momentum, steps multiplied by step_expand after every accepted iteration.  The model's
    squared gradient norms are not used."""

    def __init__(self, model, *, max_iterations=100, max_line_search_iterations=10,
                 initial_step_size=0.001, step_expand=1.2, step_shrink=0.5,
                 adaptive_tolerance=1e-4, multivariate_line_search=True, callback=None):
        self.model = model
        self.max_iterations = int(max_iterations)
        self.max_line_search_iterations = int(max_line_search_iterations)
        self.initial_step_size = float(initial_step_size)
        self.expand = float(step_expand)
        self.shrink = float(step_shrink)
        self.tolerance = float(adaptive_tolerance)
        self.multivariate = bool(multivariate_line_search)
        self.callback = callback

    def Update(self, population_rer=None, individual_rer=None):
        m = self.model
        step_of = FastGradientAscent._step
        evaluate = lambda f, p, i: _evaluate(m, f, p, i)  # noqa: E731
        fixed = _copy(m.GetFixedEffects())
        pop_rer = _copy(population_rer or {})
        ind_rer = _copy(individual_rer or {})
        history = [np.zeros(0) for _ in range(self.max_iterations + 1)]
        terms, _ = m.UpdateFixedEffectsAndComputeCompleteLogLikelihood(pop_rer, ind_rer)
        history[0] = np.asarray(terms, dtype=np.float64)
        lsq_ref = float(history[0].sum())
        pop_grad, ind_grad, _ = m.ComputeCompleteLogLikelihoodGradient(pop_rer, ind_rer)
        nb_params = len(pop_grad) + len(ind_grad)
        step = np.full(nb_params, self.initial_step_size)
        it = 1
        while it < self.max_iterations + 1:  # GradientAscent.cxx:112-220
            found = False
            for _li in range(self.max_line_search_iterations):
                new_fixed, new_pop, new_ind = step_of(fixed, pop_rer, ind_rer, pop_grad, ind_grad,
                                                      step)
                new_terms, oob = evaluate(new_fixed, new_pop, new_ind)
                if float(np.sum(new_terms)) - lsq_ref >= 0 and not oob:
                    found = True
                    break
                if nb_params > 1 and self.multivariate:
                    step = step * self.shrink
                    props = []
                    qs = np.zeros(nb_params)
                    for k in range(nb_params):
                        local = step.copy()
                        local[k] /= self.shrink
                        prop = step_of(fixed, pop_rer, ind_rer, pop_grad, ind_grad, local)
                        t_k, oob_k = evaluate(*prop)
                        props.append((prop, t_k, oob_k))
                        qs[k] = float(np.sum(t_k)) - lsq_ref
                    idx = int(np.argmax(qs))
                    if qs[idx] >= 0 and not props[idx][2]:
                        (new_fixed, new_pop, new_ind), new_terms, _ = props[idx]
                        step[idx] /= self.shrink
                        found = True
                        break
                else:
                    step = step * self.shrink
            if not found:
                m.SetFixedEffects(fixed)  # line search exhausted (cxx:177-182)
                break
            history[it] = new_terms
            fixed, pop_rer, ind_rer = new_fixed, new_pop, new_ind
            step = step * self.expand
            lsq_ref = float(np.sum(new_terms))
            if self.callback:
                self.callback(it, history[it])
            d_cur = history[it - 1].sum() - history[it].sum()
            d_ref = history[0].sum() - history[it].sum()
            if abs(d_cur) < self.tolerance * abs(d_ref):
                break  # tolerance met before the gradient update (cxx:199-203)
            m.SetFixedEffects(fixed)
            pop_grad, ind_grad, _ = m.ComputeCompleteLogLikelihoodGradient(pop_rer, ind_rer)
            it += 1
        return {
            "computation_end_state": True, "iter": it, "step": step,
            "log_likelihood_history": history, "fixed_effects": fixed,
            "population_rer": pop_rer, "individual_rer": ind_rer, "pop_grad": pop_grad,
            "ind_grad": ind_grad,
        }
