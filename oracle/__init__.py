"""FP64 CPU oracle for the Markov-GP inference hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``bayesnewton_amd`` imports, links or runs
anything from this package: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker.

It is a float64 NumPy restatement of the reference algorithm (AaltoML/BayesNewton
v1.3.4).  Every function cites the reference ``file:line`` it follows.  The
reference itself (JAX 0.4.14 + objax 1.7.0) is not importable in this image, so
the oracle is pinned by closed forms instead (see ``tests/test_oracle_*.py``):

* ``kernels``     - closed-form A(dt) vs ``scipy.linalg.expm(F dt)``; Pinf vs the
                    continuous Lyapunov equation; Q vs the Van Loan integral;
                    H expm(F|tau|) Pinf H' vs the covariance function K(tau).
* ``ops``         - filter log-likelihood vs the dense log N(y | 0, K + R);
                    smoother moments vs the dense GP posterior; sequential vs
                    associative-scan forms.  Mean-field forms vs the dense filter
                    (diagonal noise), the exact first-step conditional and the
                    decoupled scalar-filter identity; infinite-horizon forms vs
                    scipy's DARE and the exact steady state; the pairs filter vs
                    dense Gaussian conditioning of the state chain.
* ``likelihoods`` - Gauss-Hermite vs ``scipy.integrate.quad``; closed forms
                    (Gaussian, Poisson, probit moment matching); analytic
                    derivatives vs central finite differences.
* ``models``      - Gaussian + VI + lr=1 gives the exact marginal likelihood;
                    Markov GP == dense GP for every inference scheme; prediction
                    vs the dense GP predictive (temporal and space-time); the
                    mean-field model == the dense one when space decouples; the
                    infinite-horizon prediction == the dense one in steady state.
* ``kfrts.c``     - the FP64 C sequential filter/smoother (CPU baseline) vs the
                    NumPy restatement.

Functions with no such check say "parity unpinned" in their docstring.
"""
