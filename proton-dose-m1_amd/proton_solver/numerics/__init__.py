"""Numerics of the M1 energy march; the march itself runs on the MI355X."""
