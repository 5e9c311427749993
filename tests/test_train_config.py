"""Host-side training configuration: the occupancy threshold cap."""
import math

from jaxngp_amd.train import Trainer, TrainConfig, density_threshold_from_min_step_size


def test_threshold_cap_follows_the_reference_formula():
    """types.py:1227-1239: 0.01 * diagonal_n_steps / (2 min(bound, 1) sqrt 3) — the
    density at which one smallest step has opacity 1 - exp(-0.01) ~ 1%."""
    cap = Trainer.threshold_cap_of(TrainConfig(), 1.0)
    assert cap == density_threshold_from_min_step_size(1024, 1.0)
    assert math.isclose(cap, 0.01 * 1024 / (2 * math.sqrt(3)), rel_tol=1e-12)
    ds = 2 * math.sqrt(3) / 1024  # the smallest step at bound 1
    assert math.isclose(1 - math.exp(-cap * ds), 1 - math.exp(-0.01), rel_tol=1e-12)
    # the formula takes min(bound, 1): bounds past 1 keep the bound-1 cap
    assert Trainer.threshold_cap_of(TrainConfig(), 4.0) == cap


def test_guided_cap_uses_the_guided_sample_count():
    c = TrainConfig(precision="fp16", proposal=True, proposal_guided=True)
    n = c.proposal_samples_per_ray + c.proposal_uniform_samples
    assert n == 64
    assert Trainer.threshold_cap_of(c, 1.0) == density_threshold_from_min_step_size(n, 1.0)
    assert math.isclose(Trainer.threshold_cap_of(c, 1.0) * 16, Trainer.threshold_cap_of(TrainConfig(), 1.0))


def test_explicit_cap_wins():
    assert Trainer.threshold_cap_of(TrainConfig(density_threshold=0.5), 1.0) == 0.5
    c = TrainConfig(precision="fp16", proposal=True, proposal_guided=True, density_threshold=2.0)
    assert Trainer.threshold_cap_of(c, 1.0) == 2.0


def test_proposal_render_settings_follow_how_the_proposal_trained():
    from jaxngp_amd import _lib
    guided = Trainer.proposal_render_settings_of(TrainConfig(precision="fp16", proposal=True, proposal_guided=True))
    assert guided == {"n_per_ray": 32, "mass_threshold": 0.5, "placement": _lib.PROPOSAL_SUPPORT}
    beside = Trainer.proposal_render_settings_of(TrainConfig(precision="fp16", proposal=True))
    assert beside == {"n_per_ray": 48, "mass_threshold": 0.98, "placement": _lib.PROPOSAL_QUANTILES}
    # the marcher-trained settings are render_image_proposal's own defaults
    import inspect
    from jaxngp_amd.renderer import render_image_proposal
    sig = inspect.signature(render_image_proposal).parameters
    assert {k: sig[k].default for k in beside} == beside
