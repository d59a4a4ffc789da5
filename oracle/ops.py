"""FP64 Kalman filter / RTS smoother of the oracle (test infrastructure only).

Restates bayesnewton/ops.py:149-380 of the reference: the sequential
(lax.scan) and associative-scan forms of the filter and the smoother;
ops.py:429-706, their mean-field forms for space-time kernels (block-diagonal
state covariance); ops.py:796-1068, the infinite-horizon (Riccati steady-state)
forms; and ops.py:383-426, the filter over pairs of states.
Shapes follow the reference: y [N,p,1], noise_cov [N,p,p], means [N,d,1],
covs [N,d,d].
"""
import numpy as np

from .utils import mvn_logpdf, solve, inv


def T(P):
    return np.swapaxes(P, -1, -2)


def process_noise_covariance(A, Pinf):
    """ops.py:149-151: Q = Pinf - A Pinf A'."""
    return Pinf - A @ Pinf @ T(A)


def discretise(kernel, dt):
    """ops.py:274-279: As = vmap(state_transition)(dt), Qs = Pinf - A Pinf A'."""
    Pinf = kernel.stationary_covariance()
    As = kernel.state_transition_batch(dt)
    Qs = process_noise_covariance(As, Pinf)
    return Pinf, As, Qs


def _sequential_kf(As, Qs, H, ys, noise_covs, m0, P0, masks, return_predict=False):
    """ops.py:154-180"""
    N = As.shape[0]
    d = As.shape[1]
    m, P, ell = m0, P0, 0.0
    out_m = np.empty((N, d, 1))
    out_P = np.empty((N, d, d))
    for n in range(N):
        A, Q = As[n], Qs[n]
        m_ = A @ m
        P_ = A @ P @ A.T + Q
        obs_mean = H @ m_
        HP = H @ P_
        S = HP @ H.T + noise_covs[n]
        ell += float(mvn_logpdf(ys[n], obs_mean, S, masks[n]))
        K = solve(S, HP).T
        m = m_ + K @ (ys[n] - obs_mean)
        P = P_ - K @ HP
        if return_predict:
            out_m[n], out_P[n] = m_, P_
        else:
            out_m[n], out_P[n] = m, P
    return ell, out_m, out_P


def parallel_filtering_element(A, Q, H, noise_cov, y):
    """ops.py:183-200 (vectorised over the leading axis)."""
    HQ, HA = H @ Q, H @ A
    S = HQ @ H.T + noise_cov
    SinvH = solve(S, np.broadcast_to(H, S.shape[:-2] + H.shape))
    K = Q @ T(SinvH)
    AA = A - K @ HA
    b = K @ y
    C = Q - K @ HQ
    SinvHA = T(SinvH @ A)
    eta = SinvHA @ y
    J = SinvHA @ HA
    return AA, b, C, J, eta


def parallel_filtering_operator(elem1, elem2):
    """ops.py:203-219 (vectorised over the leading axis)."""
    A1, b1, C1, J1, eta1 = elem1
    A2, b2, C2, J2, eta2 = elem2
    d = A1.shape[-1]
    C1inv = inv(C1)
    temp = np.linalg.solve(C1inv + J2, C1inv)
    A2temp = A2 @ temp
    AA = A2temp @ A1
    b = A2temp @ (b1 + C1 @ eta2) + b2
    C = A2temp @ C1 @ T(A2) + C2
    A1temp = T(A1) @ T(temp)
    eta = A1temp @ (eta2 - J2 @ b1) + eta1
    J = A1temp @ J2 @ A1 + J1
    return AA, b, C, J, eta


def make_associative_filtering_elements(As, Qs, H, ys, noise_covs, m0, P0):
    """ops.py:222-229"""
    Qs = Qs.copy()
    Qs[0] = P0
    AA, b, C, J, eta = parallel_filtering_element(As, Qs, H, noise_covs, ys)
    S = H @ Qs[0] @ H.T + noise_covs[0]
    K0 = solve(S, H @ Qs[0]).T
    b = b.copy()
    b[0] = b[0] + m0 - K0 @ H @ m0
    return AA, b, C, J, eta


def associative_scan(op, elems, reverse=False):
    """Inclusive scan with an associative operator (Hillis-Steele, vectorised).

    Stands in for jax.lax.associative_scan (ops.py:6): the same prefix results
    in exact arithmetic (the combine order differs, so results agree to
    rounding).  op(a, b) combines a (earlier) with b (later), as in the
    reference's filtering operator; with reverse=True the scan runs from the
    end and op(a, b) receives the later element first, as in ops.py:348.
    """
    elems = [np.array(e, dtype=np.float64) for e in elems]
    N = elems[0].shape[0]
    if reverse:
        elems = [e[::-1].copy() for e in elems]
    shift = 1
    while shift < N:
        prev = [e[:-shift] for e in elems]
        cur = [e[shift:] for e in elems]
        comb = op(prev, cur)
        elems = [np.concatenate([e[:shift], c], axis=0) for e, c in zip(elems, comb)]
        shift *= 2
    if reverse:
        elems = [e[::-1].copy() for e in elems]
    return elems


def _parallel_kf(As, Qs, H, ys, noise_covs, m0, P0, masks, return_predict=False):
    """ops.py:237-253"""
    init = make_associative_filtering_elements(As, Qs, H, ys, noise_covs, m0, P0)
    final = associative_scan(parallel_filtering_operator, init)
    fms, fPs = final[1], final[2]
    mpredict = As @ np.concatenate([m0[None], fms[:-1]])
    Ppredict = As @ np.concatenate([P0[None], fPs[:-1]]) @ T(As) + Qs
    loglik = float(np.sum(mvn_logpdf(ys, H @ mpredict, H @ Ppredict @ H.T + noise_covs, masks)))
    if return_predict:
        return loglik, mpredict, Ppredict
    return loglik, fms, fPs


def kalman_filter(dt, kernel, y, noise_cov, mask=None, parallel=False, return_predict=False):
    """ops.py:256-285"""
    y = np.asarray(y, dtype=np.float64)
    noise_cov = np.asarray(noise_cov, dtype=np.float64)
    if mask is None:
        mask = np.zeros_like(y, dtype=bool)
    mask = np.asarray(mask, dtype=bool).reshape(y.shape)
    Pinf, As, Qs = discretise(kernel, dt)
    minf = np.zeros([Pinf.shape[0], 1])
    H = kernel.measurement_model()
    f = _parallel_kf if parallel else _sequential_kf
    ell, means, covs = f(As, Qs, H, y, noise_cov, minf, Pinf, mask, return_predict=return_predict)
    return ell, (means, covs)


def _sequential_rts(fms, fPs, As, Qs, H, return_full):
    """ops.py:288-311"""
    N, d = fms.shape[0], fms.shape[1]
    sm, sP = fms[-1], fPs[-1]
    sms = np.empty((N,) + (fms.shape[1:] if return_full else (H.shape[0], 1)))
    sPs = np.empty((N,) + (fPs.shape[1:] if return_full else (H.shape[0], H.shape[0])))
    gains = np.empty((N, d, d))
    for n in range(N - 1, -1, -1):
        fm, fP, A, Q = fms[n], fPs[n], As[n], Qs[n]
        pm = A @ fm
        AfP = A @ fP
        pP = AfP @ A.T + Q
        C = solve(pP, AfP).T
        sm = fm + C @ (sm - pm)
        sP = fP + C @ (sP - pP) @ C.T
        gains[n] = C
        if return_full:
            sms[n], sPs[n] = sm, sP
        else:
            sms[n], sPs[n] = H @ sm, H @ sP @ H.T
    return sms, sPs, gains


def parallel_smoothing_element(A, Q, m, P):
    """ops.py:318-325"""
    Pp = A @ P @ T(A) + Q
    E = T(solve(Pp, A @ P))
    g = m - E @ A @ m
    L = P - E @ Pp @ T(E)
    return E, g, L


def parallel_smoothing_operator(elem1, elem2):
    """ops.py:328-335 - elem1 is the later element in time (reverse scan)."""
    E1, g1, L1 = elem1
    E2, g2, L2 = elem2
    return E2 @ E1, E2 @ g1 + g2, E2 @ L1 @ T(E2) + L2


def _parallel_rts(fms, fPs, As, Qs, H, return_full):
    """ops.py:338-354"""
    E, g, L = parallel_smoothing_element(As, Qs, fms, fPs)
    gains = E.copy()
    E = E.copy(); g = g.copy(); L = L.copy()
    E[-1] = 0.0
    g[-1] = fms[-1]
    L[-1] = fPs[-1]
    final = associative_scan(parallel_smoothing_operator, [E, g, L], reverse=True)
    sms, sPs = final[1], final[2]
    if return_full:
        return sms, sPs, gains
    return H @ sms, H @ sPs @ H.T, gains


def rauch_tung_striebel_smoother(dt, kernel, filter_mean, filter_cov, return_full=False, parallel=False):
    """ops.py:357-380 (dt is the already-shifted step vector)."""
    Pinf, As, Qs = discretise(kernel, dt)
    H = kernel.measurement_model()
    f = _parallel_rts if parallel else _sequential_rts
    return f(np.asarray(filter_mean, np.float64), np.asarray(filter_cov, np.float64), As, Qs, H, return_full)


# ---- mean-field space-time filter / smoother (ops.py:429-706) -------------------------
# The state is M independent temporal blocks of size d (SpatioTemporalKernel's
# *_meanfield forms, kernels.py:526-575); after every update only the diagonal
# blocks of the state covariance are kept.  Written with full matrices, as the
# reference is, so that the GPU's block-reduced formulas are checked against it.

def block_diag(Pb):
    """[..., M, d, d] blocks -> [..., Md, Md] (build_block_diag, ops.py:431-433)."""
    M, d = Pb.shape[-3], Pb.shape[-1]
    out = np.zeros(Pb.shape[:-3] + (M * d, M * d))
    for i in range(M):
        out[..., i * d:(i + 1) * d, i * d:(i + 1) * d] = Pb[..., i, :, :]
    return out


def diag_blocks(P, M, d):
    """[..., Md, Md] -> its diagonal blocks [..., M, d, d] (get_block_cov, ops.py:438-439)."""
    return np.stack([P[..., i * d:(i + 1) * d, i * d:(i + 1) * d] for i in range(M)], axis=-3)


def discretise_meanfield(kernel, dt):
    """ops.py:599-603: Pinf [M,d,d] (stationary_covariance_meanfield, kernels.py:526-532),
    As, Qs [N,M,d,d] (state_transition_meanfield, kernels.py:553-561)."""
    kt, M = kernel.temporal_kernel, kernel.M
    Pt = kt.stationary_covariance()
    At = kt.state_transition_batch(dt)
    Qt = process_noise_covariance(At, Pt)
    N, d = At.shape[0], At.shape[1]
    Pinf = np.broadcast_to(Pt, (M, d, d)).copy()
    As = np.broadcast_to(At[:, None], (N, M, d, d)).copy()
    Qs = np.broadcast_to(Qt[:, None], (N, M, d, d)).copy()
    return Pinf, As, Qs


def _sequential_kf_mf(As, Qs, H, ys, noise_covs, m0, P0, masks):
    """ops.py:429-465"""
    N, M, d = As.shape[0], As.shape[1], As.shape[2]
    m, P, ell = m0, P0, 0.0
    fms = np.empty((N, M, d, 1))
    fPs = np.empty((N, M, d, d))
    for n in range(N):
        mf = (As[n] @ m).reshape(-1, 1)
        Pf = block_diag(As[n] @ P @ T(As[n]) + Qs[n])
        obs_mean = H @ mf
        HP = H @ Pf
        S = HP @ H.T + noise_covs[n]
        ell += float(mvn_logpdf(ys[n], obs_mean, S, masks[n]))
        K = solve(S, HP).T
        m = (mf + K @ (ys[n] - obs_mean)).reshape(M, d, 1)
        P = diag_blocks(Pf - K @ HP, M, d)
        fms[n], fPs[n] = m, P
    return ell, fms, fPs


def parallel_filtering_element_mf(A, Q, H, noise_cov, y):
    """ops.py:468-496 (vectorised over the leading time axis; A, Q [N,M,d,d])."""
    M, d = A.shape[-3], A.shape[-1]
    Qf, Af = block_diag(Q), block_diag(A)
    HQ, HA = H @ Qf, H @ Af
    S = HQ @ H.T + noise_cov
    SinvH = solve(S, np.broadcast_to(H, S.shape[:-2] + H.shape))
    K = Qf @ T(SinvH)
    AA = diag_blocks(Af - K @ HA, M, d)
    b = (K @ y).reshape(y.shape[:-2] + (M, d, 1))
    C = diag_blocks(Qf - K @ HQ, M, d)
    SinvHA = T(SinvH @ Af)
    eta = (SinvHA @ y).reshape(y.shape[:-2] + (M, d, 1))
    J = diag_blocks(SinvHA @ HA, M, d)
    return AA, b, C, J, eta


def _parallel_kf_mf(As, Qs, H, ys, noise_covs, m0, P0, masks):
    """ops.py:523-578: elements (with Q_0 = P0 and the m0 correction of b_0), the
    blockwise associative scan, then the log likelihood from the predictions."""
    N, M, d = As.shape[0], As.shape[1], As.shape[2]
    Qs0 = Qs.copy()
    Qs0[0] = P0
    AA, b, C, J, eta = parallel_filtering_element_mf(As, Qs0, H, noise_covs, ys)
    Q0 = block_diag(P0)
    S = H @ Q0 @ H.T + noise_covs[0]
    K0 = solve(S, H @ Q0).T
    mv = m0.reshape(-1, 1)
    b = b.copy()
    b[0] = b[0] + (mv - K0 @ (H @ mv)).reshape(M, d, 1)
    final = associative_scan(parallel_filtering_operator, [AA, b, C, J, eta])
    fms, fPs = final[1], final[2]
    mpred = (As @ np.concatenate([m0[None], fms[:-1]])).reshape(N, M * d, 1)
    Ppred = block_diag(As @ np.concatenate([P0[None], fPs[:-1]]) @ T(As) + Qs)
    loglik = float(np.sum(mvn_logpdf(ys, H @ mpred, H @ Ppred @ H.T + noise_covs, masks)))
    return loglik, fms, fPs


def kalman_filter_meanfield(dt, kernel, y, noise_cov, mask=None, parallel=False):
    """ops.py:581-608: returns ell, (means [N,M,d,1], covs [N,M,d,d])."""
    y = np.asarray(y, dtype=np.float64)
    noise_cov = np.asarray(noise_cov, dtype=np.float64)
    if mask is None:
        mask = np.zeros_like(y, dtype=bool)
    mask = np.asarray(mask, dtype=bool).reshape(y.shape)
    Pinf, As, Qs = discretise_meanfield(kernel, dt)
    minf = np.zeros(Pinf.shape[:2] + (1,))
    H = kernel.measurement_model()
    f = _parallel_kf_mf if parallel else _sequential_kf_mf
    ell, means, covs = f(As, Qs, H, y, noise_cov, minf, Pinf, mask)
    return ell, (means, covs)


def _sequential_rts_mf(fms, fPs, As, Qs, H, return_full):
    """ops.py:611-641"""
    N = fms.shape[0]
    sm, sP = fms[-1], fPs[-1]
    sms, sPs, gains = [None] * N, [None] * N, [None] * N
    for n in range(N - 1, -1, -1):
        fm, fP, A, Q = fms[n], fPs[n], As[n], Qs[n]
        pm = A @ fm
        AfP = A @ fP
        pP = AfP @ T(A) + Q
        C = T(solve(pP, AfP))
        sm = fm + C @ (sm - pm)
        sP = fP + C @ (sP - pP) @ T(C)
        gains[n] = block_diag(C)
        if return_full:
            sms[n], sPs[n] = sm.reshape(-1, 1), block_diag(sP)
        else:
            sms[n], sPs[n] = H @ sm.reshape(-1, 1), H @ block_diag(sP) @ H.T
    return np.stack(sms), np.stack(sPs), np.stack(gains)


def _parallel_rts_mf(fms, fPs, As, Qs, H, return_full):
    """ops.py:644-678"""
    N, M, d = fms.shape[0], fms.shape[1], fms.shape[2]
    E, g, L = parallel_smoothing_element(As, Qs, fms, fPs)
    gains = block_diag(E)
    E = E.copy(); g = g.copy(); L = L.copy()
    E[-1] = 0.0
    g[-1] = fms[-1]
    L[-1] = fPs[-1]
    final = associative_scan(parallel_smoothing_operator, [E, g, L], reverse=True)
    sms = final[1].reshape(N, M * d, 1)
    sPs = block_diag(final[2])
    if return_full:
        return sms, sPs, gains
    return H @ sms, H @ sPs @ H.T, gains


def rauch_tung_striebel_smoother_meanfield(dt, kernel, filter_mean, filter_cov, return_full=False,
                                           parallel=False):
    """ops.py:681-706 (dt is the already-shifted step vector)."""
    Pinf, As, Qs = discretise_meanfield(kernel, dt)
    H = kernel.measurement_model()
    f = _parallel_rts_mf if parallel else _sequential_rts_mf
    return f(np.asarray(filter_mean, np.float64), np.asarray(filter_cov, np.float64), As, Qs, H, return_full)


# ---- infinite-horizon filter / smoother (ops.py:796-1068) ----------------------------
# A fixed step (dt[1]) and a steady-state covariance from a fixed number of Riccati
# iterations started at `dare_init`; only the means run over the data.

def dare(A, H, Q, R, Pinit, num_iters=20):
    """ops.py:796-823: X <- A (X - K H X) A' + Q, K = (S^-1 H X)', S = H X H' + R."""
    X = np.array(Pinit, dtype=np.float64)
    for _ in range(num_iters):
        HX = H @ X
        S = HX @ H.T + R
        K = solve(S, HX).T
        X = A @ (X - K @ HX) @ A.T + Q
    return X


def _ih_affine_scan(Ms, vs):
    """Prefix compositions x_n = M_n x_{n-1} + v_n from x = 0 (the maps of
    ops.py:849-874 composed in a tree: (M1, v1) then (M2, v2) -> (M2 M1, M2 v1 + v2))."""
    def op(e1, e2):
        M1, v1 = e1
        M2, v2 = e2
        return M2 @ M1, M2 @ v1 + v2
    return associative_scan(op, [Ms, vs])[1]


def kalman_filter_infinite_horizon(dt, kernel, y, noise_cov, mask=None, parallel=False, heteroscedastic=False,
                                   noise_cov_tied=None, dare_iters=20, dare_init=None):
    """ops.py:877-953: returns ell, (means [N,d,1], (Pdare, cov)).  The associative form
    composes the same affine maps (ops.py:849-874; its heteroscedastic elements are their
    inverses), so it is restated by the scan of the maps themselves."""
    y = np.asarray(y, dtype=np.float64)
    noise_cov = np.asarray(noise_cov, dtype=np.float64)
    if mask is None:
        mask = np.zeros_like(y, dtype=bool)
    mask = np.asarray(mask, dtype=bool).reshape(y.shape)
    dt = np.asarray(dt, dtype=np.float64).reshape(-1)
    Pinf = kernel.stationary_covariance()
    m0 = np.zeros([Pinf.shape[0], 1])
    A = kernel.state_transition(dt[1])
    Q = process_noise_covariance(A, Pinf)
    H = kernel.measurement_model()
    Pdare = dare(A, H, Q, noise_cov_tied, Pinf if dare_init is None else dare_init, dare_iters)
    S = H @ Pdare @ H.T + noise_cov_tied
    K = Pdare @ solve(S, H).T
    HA = H @ A
    N = dt.shape[0]
    if heteroscedastic:
        Ss = H @ Pdare @ H.T + noise_cov
        Ks = Pdare @ T(solve(Ss, np.broadcast_to(H, (N,) + H.shape)))
        AKHAs = A - Ks @ HA
    else:
        Ss = np.broadcast_to(S, (N,) + S.shape)
        Ks = np.broadcast_to(K, (N,) + K.shape)
        AKHAs = np.broadcast_to(A - K @ HA, (N,) + A.shape)
    Kys = Ks @ y
    cov = Pdare - K @ H @ Pdare
    if parallel:
        means = _ih_affine_scan(np.array(AKHAs), np.array(Kys))
    else:
        means = np.empty((N,) + m0.shape)
        m = m0
        for n in range(N):
            m = AKHAs[n] @ m + Kys[n]
            means[n] = m
    obs_means = HA @ np.concatenate([m0[None], means[:-1]])
    ell = float(np.sum(mvn_logpdf(y, obs_means, np.array(Ss), mask)))
    return ell, (means, (Pdare, cov))


def rts_dare(A, Q, Pinf, num_iters=20):
    """ops.py:956-973: X <- A X A' + Q."""
    X = np.array(Pinf, dtype=np.float64)
    for _ in range(num_iters):
        X = A @ X @ A.T + Q
    return X


def rauch_tung_striebel_smoother_infinite_horizon(dt, kernel, filter_mean, filter_cov, return_full=False,
                                                  parallel=False, dare_iters=20, dare_init=None):
    """ops.py:1018-1068 (dt is the smoother's step vector; A from its first entry):
    returns means, covs, gains, dare_cov.  Both forms run the same constant-gain
    recursion, sm_n = fm_n + G (sm_{n+1} - A fm_n) from sm = fm_{N-1}."""
    dt = np.asarray(dt, dtype=np.float64).reshape(-1)
    Pinf = kernel.stationary_covariance()
    A = kernel.state_transition(dt[0])
    H = kernel.measurement_model()
    N = dt.shape[0]
    Pdare, fcov = filter_cov
    gain = fcov @ solve(Pdare, A).T
    Qdare = fcov - gain @ Pdare @ gain.T
    dare_cov = rts_dare(gain, Qdare, Pinf if dare_init is None else dare_init, dare_iters)
    fms = np.asarray(filter_mean, dtype=np.float64)
    Afms = A @ fms
    if parallel:
        # ops.py:990-1012: elements (G, fm - G A fm) and a last (0, fm_{N-1}), reverse scan
        Gs = np.concatenate([np.broadcast_to(gain, (N,) + gain.shape), np.zeros((1,) + gain.shape)])
        vs = np.concatenate([fms - gain @ Afms, fms[-1:]])
        final = associative_scan(lambda e1, e2: (e2[0] @ e1[0], e2[0] @ e1[1] + e2[1]), [Gs, vs], reverse=True)
        sms = final[1][:-1]
    else:
        sms = np.empty_like(fms)
        sm = fms[-1]
        for n in range(N - 1, -1, -1):
            sm = fms[n] + gain @ (sm - Afms[n])
            sms[n] = sm
    if return_full:
        means, cov = sms, dare_cov
    else:
        means, cov = H @ sms, H @ dare_cov @ H.T
    covs = np.broadcast_to(cov, (N,) + cov.shape).copy()
    gains = np.broadcast_to(gain, (N,) + gain.shape).copy()
    return means, covs, gains, dare_cov


# ---- Kalman filter over pairs of states (ops.py:383-426), the sparse Markov model's filter --

def kalman_filter_pairs(dt, kernel, y, noise_cov, mask=None, parallel=False):
    """ops.py:383-426: the state is the pair (x_{k-1}, x_k) of neighbouring inducing states,
    A_pair = [[0, I], [0, A]], Q_pair = blkdiag(1e-32 I, Q), observed whole (H = I).
    y [N,2d,1], noise_cov [N,2d,2d].  Returns ell, (means [N-1,d,1], covs [N-1,d,d]): the
    first state of each pair after the first step."""
    y = np.asarray(y, dtype=np.float64)
    noise_cov = np.asarray(noise_cov, dtype=np.float64)
    if mask is None:
        mask = np.zeros_like(y, dtype=bool)
    mask = np.asarray(mask, dtype=bool).reshape(y.shape)
    Pinf, As, Qs = discretise(kernel, dt)
    d = Pinf.shape[0]
    Z = np.zeros((d, d))
    Pinfpair = np.block([[Pinf, Z], [Z, Pinf]])
    minfpair = np.zeros((2 * d, 1))
    N = As.shape[0]
    Apairs = np.zeros((N, 2 * d, 2 * d))
    Qpairs = np.zeros((N, 2 * d, 2 * d))
    Apairs[:, :d, d:] = np.eye(d)
    Apairs[:, d:, d:] = As
    Qpairs[:, :d, :d] = 1e-32 * np.eye(d)
    Qpairs[:, d:, d:] = Qs
    H = np.eye(2 * d)
    f = _parallel_kf if parallel else _sequential_kf
    ell, means, covs = f(Apairs, Qpairs, H, y, noise_cov, minfpair, Pinfpair, mask)
    return ell, (means[1:, :d], covs[1:, :d, :d])
