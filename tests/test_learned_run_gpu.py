"""A short learned run beside its FP64 shadow (nrg_amd/learned_run.py, the
driver of the 512x512x256, 200-step configuration): both fields take the same
seeded advection shift after every split step, the learned field's
accumulated drift against the shadow stays within the learned step's bars at
every step, retraining happens on schedule on pairs that include the cells the
learned step sent to its FP64 fallback, and a retrained network is kept only
when the learned step on held-out states falls back less without a larger
error."""
import pytest

pytestmark = pytest.mark.gpu


def test_learned_run_shift_drift_and_fallback_retraining():
    from nrg_amd import learned_run
    rep = learned_run.run(shape=(48, 40, 24), steps=12, retrain_every=4, audit_frac=0.05,
                          retrain_epochs=1, fallback_pairs=4096, shadow_pairs=4096,
                          log=lambda m: print("[learned_run] " + m, flush=True))
    steps = rep["per_step"]
    assert rep["cells"] == 48 * 40 * 24 and len(steps) == 12
    # the field moves: a seeded shift every step, not all of them zero
    assert any(r["shift"] != [0, 0] for r in steps)
    for r in steps:
        assert r["max_rel_T"] <= rep["bars"]["rel_T"], r
        assert r["max_abs_Y_major"] <= rep["bars"]["abs_Y_major"], r
        assert r["sample_p99_rel_T"] <= r["max_rel_T"] + 1e-15
        assert 0.0 <= r["fallback_frac"] <= 1.0 and 0.0 <= r["network_frac"] <= 1.0
    assert rep["steps_within_bars"] == 12
    assert [t["after_step"] for t in rep["retrains"]] == [3, 7]
    assert sum(t["fallback_pairs"] for t in rep["retrains"]) > 0
    for t in rep["retrains"]:
        assert t["pairs"] > 0 and t["val_states"] > 0, t
        assert t["kept"] == (t["val_fallback_after"] < t["val_fallback_before"] and
                             t["val_max_abs_dY_major_after"] <= t["val_max_abs_dY_major_before"])
    assert rep["hbm_in_use_peak_GB"] > rep["field_GB"] > rep["state_GB"] > 0
