"""Run table shared by the fixture writer and the parity tests.

Each entry: key -> (entry point name, keyword builder).  The builder receives
the fixture/spec dict (``sigma``, ``sigma_map``, ``raw_tau``) and returns the
method-specific keyword arguments; patch shape, overlap, mask and mask
threshold are common to every run of a volume.
"""

RECOMBS = ("weighted", "average", "center")


def run_table(spec):
    sig = float(spec["sigma"])
    runs = []
    for recomb in RECOMBS:
        runs.append((f"mppca_ts1_{recomb}", "mp_pca",
                     dict(recombination=recomb, threshold_scale=1.0)))
    runs.append(("mppca_ts23_weighted", "mp_pca",
                 dict(recombination="weighted", threshold_scale=2.3)))
    runs.append(("hybrid_scalar_weighted", "hybrid_pca",
                 dict(noise_std=sig, recombination="weighted", progbar=False)))
    runs.append(("hybrid_map_average", "hybrid_pca",
                 dict(noise_std=spec["sigma_map"], recombination="average", progbar=False)))
    runs.append(("raw_weighted", "raw_svt",
                 dict(threshold=float(spec["raw_tau"]), recombination="weighted",
                      progbar=False)))
    runs.append(("raw_center", "raw_svt",
                 dict(threshold=float(spec["raw_tau"]), recombination="center",
                      progbar=False)))
    if bool(spec.get("optimal_ok", True)):
        for loss in ("fro", "nuc", "ope"):
            runs.append((f"optimal_{loss}_weighted", "optimal_thresholding",
                         dict(loss=loss, recombination="weighted", progbar=False)))
        runs.append(("optimal_fro_map_average", "optimal_thresholding",
                     dict(loss="fro", noise_std=spec["sigma_map"], recombination="average",
                          progbar=False)))
        runs.append(("optimal_nuc_scalar_center", "optimal_thresholding",
                     dict(loss="nuc", noise_std=sig, recombination="center", progbar=False)))
    # NORDIC draws from the global legacy stream: callers seed np.random with 0 first.
    runs.append(("nordic_seed0_weighted", "nordic",
                 dict(noise_std=sig, recombination="weighted", progbar=False)))
    return runs
