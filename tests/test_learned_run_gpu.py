"""A short learned run with the running FP64 audit and online retraining
(nrg_amd/learned_run.py, the driver of the 512x512x256, 200-step
configuration): every step's audited cells stay within the learned step's
bars against FP64 from the same state, retraining happens on schedule, and
a retrained network is kept only when the learned step on held-out audited
states falls back less without a larger error."""
import pytest

pytestmark = pytest.mark.gpu


def test_learned_run_audit_and_online_retraining():
    from nrg_amd import learned_run
    rep = learned_run.run(shape=(48, 40, 24), steps=9, retrain_every=3, audit_frac=0.05,
                          retrain_epochs=1, log=lambda m: print("[learned_run] " + m, flush=True))
    assert rep["cells"] == 48 * 40 * 24 and len(rep["per_step"]) == 9
    for r in rep["per_step"]:
        assert r["audit_max_rel_T"] <= rep["bars"]["rel_T"], r
        assert r["audit_max_abs_Y_major"] <= rep["bars"]["abs_Y_major"], r
        assert 0.0 <= r["fallback_frac"] <= 1.0 and 0.0 <= r["accepted_frac"] <= 1.0
    assert [t["after_step"] for t in rep["retrains"]] == [2, 5]
    for t in rep["retrains"]:
        # a retrained network is kept only if, on held-out audited states, the
        # learned step falls back less without a larger major-species error
        assert t["pairs"] > 0 and t["val_states"] > 0, t
        assert t["kept"] == (t["val_fallback_after"] < t["val_fallback_before"] and
                             t["val_max_abs_dY_major_after"] <= max(
                                 t["val_max_abs_dY_major_before"], 0.5 * rep["bars"]["abs_Y_major"]))
    assert rep["hbm_in_use_peak_GB"] > rep["field_GB"] > 0
