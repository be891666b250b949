"""Inputs of the adjoint tests: worlds of the throw workload, their first contacts, colours and
targets, and the float64 fixed-contact model over them (oracle/fixed_contacts.py)."""
import os
import sys

import numpy as np

from twofold_amd import scenes
from twofold_amd.world import HostWorld
import refworld

sys.path.insert(0, os.path.join(refworld.ROOT, "oracle"))
from fixed_contacts import FixedContacts  # noqa: E402


class Case:
    """`worlds` throw worlds side by side: bodies, contacts (global body indices), world starts, the
    oracle's colouring of the contacts, seeded targets and the float64 model."""

    def __init__(self, seeds, steps):
        self.steps = steps
        self.scenes = [scenes.throw_world(s, steps=steps) for s in seeds]
        self.world = HostWorld(self.scenes)
        self.settings = self.world.settings
        self.dt = float(self.world.dt)
        self.bodies = self.world.bodies()
        self.contacts, _ = self.world.collide(self.bodies, self.world.dt)
        self.body_start = np.asarray(self.world.body_start, np.uint32)
        lib = refworld._load_oracle()
        b = self.bodies.copy()
        for i in range(len(b)):
            lib.tfo_update_inverse_mass(b[i:i + 1].ctypes.data)
        self.colors = np.full(len(self.contacts), 255, np.uint8)
        lst = np.arange(len(self.contacts), dtype=np.uint32)
        lib.tfo_color(b.ctypes.data, len(b), self.contacts.ctypes.data, lst.ctypes.data, len(lst),
                      self.colors.ctypes.data)
        self.model = FixedContacts(self.bodies, self.contacts, self.colors, self.settings, self.dt)
        rng = np.random.default_rng(int(seeds[0]))
        self.target = self.model.com0 + rng.uniform(-0.5, 0.5, self.model.com0.shape)
        self.world_of_contact = np.searchsorted(self.body_start, self.contacts["body_a"], side="right") - 1

    def dl_dcom(self, com):
        g = np.zeros_like(com, dtype=np.float64)
        mv = self.model.moving
        g[mv] = 2.0 * (com[mv] - self.target[mv])
        return g

    def per_world(self, per_contact):
        return np.bincount(self.world_of_contact, weights=per_contact, minlength=len(self.scenes))
