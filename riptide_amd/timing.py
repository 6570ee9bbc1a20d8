"""Wall-clock logging decorator (timing.py:6-17 of the reference): logs the
runtime of the wrapped call at DEBUG level on the ``riptide.timing`` logger."""
import logging
import time
from functools import wraps


def timing(func):
    @wraps(func)
    def wrapped(*args, **kwargs):
        t0 = time.time()
        out = func(*args, **kwargs)
        dt = time.time() - t0
        logging.getLogger("riptide.timing").debug(
            "{!r} runtime: {:.2f} ms".format(func.__name__, dt * 1000.0))
        return out

    return wrapped
