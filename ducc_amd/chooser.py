"""Learned parameter chooser (host side).

The trained network's weights live in ducc_amd/data/chooser_weights.json;
when the file is absent the library defers to ducc0's analytic choice."""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "chooser_weights.json")


def available():
    return os.path.exists(WEIGHTS)
