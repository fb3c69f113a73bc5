"""End-to-end proposal-guided search on the GPU: the network trained by its HIP kernels on
the plain search of seeded micrographs, then the guided search on held-out micrographs
against the exhaustive search (recall of z > 7.5 peaks), beside the same search with
random bins instead of the network's (the method without the learned part) at the same
orientation budget, and with every tile sent to the exhaustive fallback (tiling and
statistics check)."""

import pytest
import torch

from leopard_em_amd import proposal

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_guided_search_recall_and_gain(cuda_device):
    out = proposal.run_experiment(cuda_device, n_train=16, n_test=3, epochs=30, log=lambda m: print(m, flush=True))
    print({k: v for k, v in out.items()}, flush=True)
    assert out["reference_peaks"] > 0
    assert out["recall_net"] is not None and out["recall_random_bins"] is not None
    assert 0.0 <= out["recall_net"] <= 1.0
    assert sum(out["tile_decisions"].values()) == 9 * 3
    assert out["loss_first_last"][1] < out["loss_first_last"][0]


@pytest.mark.timeout(600)
def test_guided_search_all_exhaustive_recovers_reference_peaks(cuda_device):
    """Every tile exhaustive (confidence threshold above 1): the tiled search with subset
    statistics finds the exhaustive full-image peaks."""
    out = proposal.run_experiment(cuda_device, n_train=2, n_test=2, epochs=1, conf=2.0, empty_conf=2.0,
                                  log=lambda m: print(m, flush=True))
    print(out, flush=True)
    assert out["tile_decisions"]["exhaustive"] == 9 * 2
    assert out["visited_fraction_net"] == 1.0
    assert out["recall_net"] >= 0.9
