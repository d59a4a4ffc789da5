"""Likelihoods of the Markov-GP hot path (mirrors bayesnewton/likelihoods.py).

Gaussian (684-803), Bernoulli with probit link (806-871), Poisson with exp
link (891-1008) and StudentsT (1011-1044).  Each holds its hyperparameters
(softplus-transformed where the reference trains them, likelihoods.py:696-706,
1019-1029) and describes itself to the HIP site kernels with `desc()`.
"""
from . import _native as nat
from .kernels import softplus, softplus_inv


class Likelihood:
    kind = None

    def desc(self):
        d = nat.LikDesc()
        d.kind = self.kind
        d.param0, d.param1 = self._params()
        return d

    def _params(self):
        return 0.0, 0.0


class Gaussian(Likelihood):
    """likelihoods.py:684-803"""
    kind = nat.BN_LIK_GAUSSIAN

    def __init__(self, variance=0.1, fix_variance=False):
        self.transformed_variance = softplus_inv(float(variance))
        self.fix_variance = fix_variance
        self.name = "Gaussian"

    @property
    def variance(self):
        return softplus(self.transformed_variance)

    @variance.setter
    def variance(self, v):
        self.transformed_variance = softplus_inv(float(v))

    def _params(self):
        return self.variance, 0.0


class Bernoulli(Likelihood):
    """likelihoods.py:806-871 (probit link: Phi(f)(1 - 2e-3) + 1e-3; y in {0, 1})."""
    kind = nat.BN_LIK_BERNOULLI_PROBIT

    def __init__(self, link="probit"):
        if link != "probit":
            raise NotImplementedError("the HIP site kernels implement the probit link")
        self.link = link
        self.name = "Bernoulli"


class Probit(Bernoulli):
    def __init__(self):
        super().__init__(link="probit")


Erf = Probit


class Poisson(Likelihood):
    """likelihoods.py:891-1008 (exp link)."""
    kind = nat.BN_LIK_POISSON

    def __init__(self, binsize=1, link="exp"):
        if link != "exp":
            raise NotImplementedError("the HIP site kernels implement the exp link")
        self.binsize = float(binsize)
        self.link = link
        self.name = "Poisson"

    def _params(self):
        return self.binsize, 0.0


class StudentsT(Likelihood):
    """likelihoods.py:1011-1044 (df fixed, scale trainable)."""
    kind = nat.BN_LIK_STUDENTS_T

    def __init__(self, scale=1.0, df=3.0, fix_scale=False):
        self.transformed_scale = softplus_inv(float(scale))
        self.df = float(df)
        self.fix_scale = fix_scale
        self.name = "Students t"

    @property
    def scale(self):
        return softplus(self.transformed_scale)

    @scale.setter
    def scale(self, v):
        self.transformed_scale = softplus_inv(float(v))

    def _params(self):
        return self.scale, self.df
