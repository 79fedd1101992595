"""Fitting the learned HH surrogate through the nsx_surrogate_* entry points."""
import time

from . import configs


def fit(engine, **overrides):
    """Generate the training transitions and train on the GPU with the recipe of
    configs.SURROGATE_RECIPE (entries overridable by keyword).  Returns the sample count, the
    per-step loss history and the wall time of each phase."""
    r = dict(configs.SURROGATE_RECIPE)
    r.update(overrides)
    t0 = time.perf_counter()
    n = engine.surrogate_generate(r["n_samples"], r["data_seed"],
                                  random_fraction=r["random_fraction"], scale_eps=r["scale_eps"])
    t1 = time.perf_counter()
    hist = engine.surrogate_train(r["steps"], waves=r["waves"],
                                  chunks_per_wave=r["chunks_per_wave"], lr=r["lr"],
                                  lr_min=r["lr_min"], seed=r["seed"], init=True)
    t2 = time.perf_counter()
    return {"samples": n, "loss": hist, "generate_s": t1 - t0, "train_s": t2 - t1, "recipe": r}
