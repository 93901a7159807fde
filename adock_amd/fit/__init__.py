"""Fitted (linear-in-weights) scoring function: term table and potential (terms), grid builder (gridfit) and the
docking driver on the kernel search (fdock) (reference: src/adock/fit)."""
