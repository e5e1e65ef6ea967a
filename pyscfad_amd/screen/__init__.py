"""Learned screen for the (T) energy: a BF16 MLP ensemble on MP2 pair features,
with the exact (T) kernel as fallback whenever the ensemble disagrees."""
