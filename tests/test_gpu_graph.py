"""The fused pass replayed from a hipGraph (MarkovGaussianProcess.use_graph) gives the
eager pass's results bit for bit: two models on the same batch run the same sequence of
inference() steps and hyper changes, one replaying its recorded pass, one launching
eagerly, with host synchronisations and eager library work between the replays; every
replay's posterior and log-likelihood equal the eager ones exactly.  The schedule tuner's
table is exercised on the same batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(seed):
    rng = np.random.default_rng(seed)
    lens = [900, 33, 2500, 1, 400]
    xs, ys = [], []
    for n in lens:
        x = np.sort(rng.uniform(0.0, n / 20.0 + 1.0, n))
        xs.append(x)
        ys.append(np.sin(x) + 0.2 * rng.standard_t(3.0, size=n))
    return xs, ys


def _model(xs, ys, graph):
    from bayesnewton_amd import kernels as K, likelihoods as L, models as M
    kern = K.Sum([K.Matern52(variance=0.8, lengthscale=1.5), K.Matern32(variance=0.4, lengthscale=4.0)])
    g = M.MarkovVariationalGP.from_series(kern, L.StudentsT(scale=0.25, df=3.0), xs, ys, chunk_len=16)
    g.use_graph = graph
    return g


def test_graph_replays_equal_eager_passes():
    import torch
    xs, ys = _batch(3)
    ge, gg = _model(xs, ys, False), _model(xs, ys, True)
    for step in range(6):
        print(f"[graph] step {step}", flush=True)
        if step == 3:  # a hyper change: the graph is recorded again
            for g in (ge, gg):
                g.kernel.components[0].lengthscale = 1.1
        ge.inference(lr=0.7)
        gg.inference(lr=0.7)
        torch.cuda.synchronize()
        for a, b in ((ge._pm, gg._pm), (ge._pv, gg._pv), (ge._loglik, gg._loglik)):
            assert torch.equal(a, b), step
        assert float(ge.energy()) == float(gg.energy())
    assert gg._graph is not None and gg._graph_key == gg._hyper_version


def test_schedule_table_roundtrip(tmp_path):
    import torch
    from bayesnewton_amd import autotune, kernels as K
    xs, ys = _batch(4)
    lens = np.array([len(x) for x in xs])
    off = np.concatenate([[0], np.cumsum(lens)])
    dev = torch.device("cuda", 0)
    desc = K.Sum([K.Matern52(), K.Matern32()]).desc()
    ls = torch.tensor([1.5, 4.0], dtype=torch.float64, device=dev)
    var = torch.tensor([0.8, 0.4], dtype=torch.float64, device=dev)
    dt = torch.as_tensor(np.concatenate([np.concatenate([[0.0], np.diff(x)]) for x in xs]), dtype=torch.float32,
                         device=dev)
    y = torch.as_tensor(np.concatenate(ys), dtype=torch.float32, device=dev)
    r = torch.full_like(y, 0.1)
    choice = autotune.tune_schedule(off, desc, ls, var, dt, y, r, device=dev, table=False, reps=2)
    assert choice["source"] == "timed" and choice["schedule"] in ("chunked", "sequential")
    path = str(tmp_path / "schedules.json")
    autotune.record_schedule(choice["signature"], {"schedule": choice["schedule"], "chunk_len": choice["chunk_len"]},
                             path=path)
    assert autotune.load_table(path)[choice["signature"]]["chunk_len"] == choice["chunk_len"]
    assert choice["signature"] == autotune.batch_signature(lens, desc, dev)
