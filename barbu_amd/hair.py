"""Python mirror of the simulation half of tcoppex/barbu's `class Hair`
(src/fx/hair.h:24-121, src/fx/hair.cc:26-125), over libftl.so.

Same names, argument meaning and error behaviour as the reference:
  init()                      hair.cc:26-40   binds the device (no shader compile)
  setup(roots, normals)       hair.cc:42-64   builds the initial state (init_simulation,
                                              :236-288, with a seeded stream) and uploads it
  update(dt)                  hair.cc:89-125  one step; a no-op on an uninitialised Hair
  set_bounding_sphere(vec4)   hair.h:72-74    collider for the next update
  initialized()               hair.h:76-78    nroots != 0
  deinit()                    hair.cc:66-87   frees the device state
The reference takes a scalp mesh resource id in setup(); mesh loading is
outside this library, so setup() takes the root positions and normals the
mesh would provide.
"""
import logging

import numpy as np

from . import Context, Groom, reference_groom

log = logging.getLogger("barbu_amd.hair")


class Hair:
    class Parameters:
        """Hair::Parameters_t (hair.h:27-51), the fields the simulation reads."""

        def __init__(self):
            self.sim_maxlength = 0.50       # sim.maxlength
            self.render_lengthScale = 1.450  # render.lengthScale -> uScaleFactor
            self.readonly_nroots = 0
            self.readonly_nControlPoints = 4  # HAIR_MAX_PARTICLE_PER_STRAND (interop.h:8)

    def __init__(self):
        self.params = Hair.Parameters()
        self._ctx = None
        self._groom = None
        self._nroots = 0
        self._bsphere = np.array([0.0, 0.0, 0.0, 1.0], np.float32)  # uBoundingSphere default

    def init(self, device=0):
        self._ctx = Context(device)

    def setup(self, roots, normals, npoints=4, seed=0):
        roots = np.asarray(roots, np.float32).reshape(-1, 3)
        if roots.shape[0] == 0:
            # hair.cc:45-48: a missing scalp logs an error and leaves nroots = 0
            log.error("The scalp mesh resource was not found.")
            return
        pos, vel = reference_groom(roots, normals, npoints=npoints, maxlength=self.params.sim_maxlength, seed=seed)
        self._groom = Groom(self._ctx, pos, vel)
        self._nroots = roots.shape[0]
        self.params.readonly_nroots = self._nroots
        self.params.readonly_nControlPoints = npoints

    def deinit(self):
        if self._groom is not None:
            self._groom.close()
            self._groom = None
        self._nroots = 0

    def set_bounding_sphere(self, bsphere):
        self._bsphere = np.asarray(bsphere, np.float32).reshape(4).copy()

    def initialized(self):
        return self._nroots != 0

    def update(self, dt, stream=None):
        if not self.initialized():
            # hair.cc:90-93
            log.debug("Calling Hair::update without initialization.")
            return
        p = self._groom.get_params()
        p.length_scale = float(self.params.render_lengthScale)
        self._groom.set_params(p)
        self._groom.set_colliders(self._bsphere)
        self._groom.step(dt, stream)

    # Accessors a renderer or a test needs (the reference's VAO reads the
    # buffers directly, hair.cc:363-391).
    def groom(self):
        return self._groom
