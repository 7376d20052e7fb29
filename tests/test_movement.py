"""Panel movement on the CPU through the jittered fake backend, whose queues
run on worker threads with random delays (test/test_movement.jl:1-106)."""
import pytest

import movement_cases as M
from fakes import JitteryBackend


def backend(rep=0):
    return JitteryBackend(jitter=2e-4, seed=rep)


@pytest.mark.parametrize("w", (20, 7, 1))
def test_synchronous_movement(w):
    M.synchronous_movement(backend(), w)


def test_movement_rejects_mismatched_buffers():
    M.movement_rejects_mismatched_buffers(backend())


def test_movement_with_narrowed_storage():
    M.movement_with_narrowed_storage(backend())


def test_queued_movement_under_jitter():
    for rep in range(20):
        M.queued_movement(backend(rep))


def test_queued_movement_matches_synchronous_movement():
    for rep in range(20):
        M.queued_matches_synchronous(backend(rep))
