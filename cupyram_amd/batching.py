"""Prefetching of CPU-side batch preparation (same interface as
cupyram/batching.py:5-152).

``AsyncBatcher(input_data_iterator, worker_func, max_prefetch=1,
max_workers=2)`` is an iterable: it keeps ``max_prefetch`` calls of
``worker_func(*args)`` in flight on a process pool and yields their results
in submission order, so CPU preparation of batch k+1 overlaps the GPU march of
batch k.  ``share_object`` / ``load_shared_object`` / ``cleanup_shared_object``
place one pickled object in named shared memory for the workers.
"""
from __future__ import annotations

import collections
import pickle
from concurrent.futures import ProcessPoolExecutor
from multiprocessing import shared_memory


class AsyncBatcher:
    """FIFO prefetching iterator over ``worker_func`` results."""

    def __init__(self, input_data_iterator, worker_func, max_prefetch=1, max_workers=2):
        self.input_iter = iter(input_data_iterator)
        self.worker_func = worker_func
        self.max_prefetch = max(1, int(max_prefetch))
        self.max_workers = max(1, int(max_workers))
        self._pending = collections.deque()
        self._executor = None

    def _submit_next(self):
        try:
            args = next(self.input_iter)
        except StopIteration:
            return False
        self._pending.append(self._executor.submit(self.worker_func, *args))
        return True

    def __iter__(self):
        with ProcessPoolExecutor(max_workers=self.max_workers) as pool:
            self._executor = pool
            for _ in range(self.max_prefetch):
                if not self._submit_next():
                    break
            while self._pending:
                result = self._pending.popleft().result()
                self._submit_next()
                yield result
        self._executor = None

    @staticmethod
    def share_object(obj):
        """Pickle ``obj`` into a new shared-memory block; returns (name, size)."""
        data = pickle.dumps(obj)
        shm = shared_memory.SharedMemory(create=True, size=max(1, len(data)))
        shm.buf[:len(data)] = data
        name = shm.name
        shm.close()
        return name, len(data)

    @staticmethod
    def load_shared_object(shm_name, size):
        """Unpickle an object written by ``share_object`` (in any process)."""
        shm = shared_memory.SharedMemory(name=shm_name)
        try:
            data = bytes(shm.buf[:size])
        finally:
            shm.close()
        return pickle.loads(data)

    @staticmethod
    def cleanup_shared_object(shm_name):
        """Unlink the block; a missing block is ignored."""
        try:
            shm = shared_memory.SharedMemory(name=shm_name)
        except FileNotFoundError:
            return
        shm.close()
        shm.unlink()
