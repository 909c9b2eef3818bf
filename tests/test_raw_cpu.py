"""The "raw" record (host code, no GPU): one profile per readout, samples x
coils data, nominal trajectory from get_kspace, header fields; reference
finalize_simulation! (src/simulation/Simulate.jl:195-214) and the shape
checks of test/unit/test_koma_compatibility.jl:99-105."""
import numpy as np

from hom_amd import get_kspace
from hom_amd.raw import RawAcquisitionData, readouts, signal_to_raw_data
from hom_amd.workloads import CONFIGS, config_sequence, make_config


def test_raw_profiles_per_readout():
    rng = np.random.Generator(np.random.PCG64(1))
    obj, seq, cfg = make_config(1, spin_subsample=5000)   # 4 interleaves x 4750 samples
    n = int(seq.adc.sum())
    sig = (rng.standard_normal((n, cfg.n_coils)) + 1j * rng.standard_normal((n, cfg.n_coils)))
    raw = signal_to_raw_data(sig[:, :, None], seq, phantom_name=obj.name,
                             sim_params={"Nblocks": 7})
    assert isinstance(raw, RawAcquisitionData)
    assert len(raw.profiles) == cfg.n_interleaves
    assert [p.data.shape for p in raw.profiles] == [(cfg.n_adc, cfg.n_coils)] * cfg.n_interleaves
    np.testing.assert_array_equal(np.vstack([p.data for p in raw.profiles]),
                                  sig.astype(np.complex64))
    traj = np.hstack([p.traj for p in raw.profiles])
    np.testing.assert_array_equal(traj, get_kspace(seq)[1].T)
    for k, p in enumerate(raw.profiles):
        assert p.head.scan_counter == k and p.head.number_of_samples == cfg.n_adc
        assert abs(p.head.sample_time_us - cfg.dwell * 1e6) < 1e-6
        assert p.traj.shape == (3, cfg.n_adc) and p.data.dtype == np.complex64
    assert raw.params["userParameters"]["Nblocks"] == 7
    assert raw.params["receiverChannels"] == cfg.n_coils


def test_readout_split_on_gaps():
    # one sequence with a single readout: one profile holding every sample
    cfg = CONFIGS[0]
    seq = config_sequence(cfg, np.random.Generator(np.random.PCG64(2)))
    groups = readouts(seq)
    assert len(groups) == 1 and groups[0].size == cfg.n_adc
