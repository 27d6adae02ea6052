"""Material images on the GPU: refused (ST_ERROR_UNSUPPORTED) at scene
creation.  The oracle restates LoadMaterial with its images
(tests/test_oracle_textures.py); compiled into the shared shading kernels the
texture path slowed the headline frame by a fifth, so it waits for kernels
instantiated for textured scenes only."""
import numpy as np
import pytest

from stratum2_amd import scenes

pytestmark = pytest.mark.gpu


def test_material_images_are_refused(gpu_ctx):
    from stratum2_amd.renderer import Scene
    with pytest.raises(RuntimeError, match="textures"):
        Scene(gpu_ctx, scenes.cornell_textured(0))
    for word in (5, 9, 12):
        sd = scenes.cornell_box(0)
        sd.images.append(np.ones((4, 4, 4), np.float32))
        w = scenes.material_record()
        w[word] = len(sd.images) - 1
        vi, n = sd.add_mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([0, 1, 2], np.uint32))
        sd.add_mesh_instance(vi, n, sd.add_material(w))
        with pytest.raises(RuntimeError):
            Scene(gpu_ctx, sd)
