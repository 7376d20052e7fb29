"""Panel movement through the HIP backend (test/test_movement.jl:1-106):
whole-panel copies with and without queues, their refusals, narrowed host
storage, and queued copies fenced by events, including a synchronous copy
issued inside submit, which runs on the submitting queue."""
import pytest

import movement_cases as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w", (20, 7, 1))
def test_synchronous_movement(backend, w):
    M.synchronous_movement(backend, w)


def test_movement_rejects_mismatched_buffers(backend):
    M.movement_rejects_mismatched_buffers(backend)


def test_movement_with_narrowed_storage(backend):
    M.movement_with_narrowed_storage(backend)


def test_queued_movement(backend):
    for _ in range(5):
        M.queued_movement(backend)


def test_queued_movement_matches_synchronous_movement(backend):
    for _ in range(5):
        M.queued_matches_synchronous(backend)
