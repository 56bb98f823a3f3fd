"""Learned cube-growth proposals (filled in by the CNN module)."""
from __future__ import annotations


class Proposer:
    @classmethod
    def load_default(cls, device):
        return None
