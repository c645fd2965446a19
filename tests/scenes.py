"""The seeded scene builders live in the package (the learned part's data
generation uses them too); tests import them from here."""
from vflip_amd.scenes import *  # noqa: F401,F403
from vflip_amd.scenes import MASK32, randhash_np, uniform01  # noqa: F401
