"""CPU oracle for the ray-traversal hot path (test infrastructure only)."""
from .oracle import *  # noqa: F401,F403
from .oracle import Scene, build, lib  # noqa: F401
