"""GPU: dvg_device_render_noalloc (the forward with its work images in the
library's static per-device scratch array) against dvg_device_render, the
entry point it mirrors, through the C ABI.

Two calls in a row on one device, the second on a smaller scene: the scratch
keeps the first call's contents, so the second call shows that the entry
point clears what must start at zero.  Images must agree within the parity
bar for images, 1e-5 absolute per channel (only the order of the float
atomics of the splat may differ).  A scene above DVG_NOALLOC_MAX_PIXELS is
rejected with InvalidScene before any launch."""
import numpy as np
import pytest

from pydvg import abi, scenes
from pydvg.scene import Scene, Shape, ShapeGroup, Circle, Solid, RenderOptions

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-5


def _render_both(renderer, sc, opts):
    import torch
    from pydvg.renderer import DeviceScene, PreparedScene
    ds = DeviceScene(renderer, PreparedScene(sc, opts))
    n = sc.width * sc.height * 4
    ref = torch.full((n,), -1.0, dtype=torch.float32, device="cuda")
    got = torch.full((n,), -2.0, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()  # the context's stream does not wait for torch's
    ds.render(ref)
    renderer.sync()
    ds.render_noalloc(got)
    renderer.sync()
    return got.cpu().numpy(), ref.cpu().numpy()


def test_noalloc_render_matches_device_render(renderer):
    cases = [
        # the largest scene the scratch holds: 1024^2, 1000 cubic paths, 2x2 spp
        ("paths1024", scenes.headline()),
        # then a smaller one with a tent filter (a splat halo) and mixed shapes
        ("mixed200x150", (scenes.mixed_scene(200, 150, 60, seed=21), scenes.config2()[1])),
    ]
    for name, (sc, opts) in cases:
        assert sc.width * sc.height <= abi.DVG_NOALLOC_MAX_PIXELS
        print(f"case noalloc {name} {sc.width}x{sc.height}", flush=True)
        got, ref = _render_both(renderer, sc, opts)
        err = float(np.abs(got - ref).max())
        print(f"{name}: max |noalloc - device_render| = {err:.3g}", flush=True)
        assert np.all(np.isfinite(got))
        assert err <= IMG_TOL


def test_noalloc_rejects_oversized_scene(renderer):
    import torch
    from pydvg.renderer import DeviceScene, PreparedScene, RenderError
    sc = Scene(1025, 1024, shapes=[Shape(Circle((100.0, 100.0), 20.0))],
               groups=[ShapeGroup(shape_indices=[0], fill=Solid((1.0, 0.0, 0.0, 1.0)))], background=(1, 1, 1, 1))
    print("case noalloc oversized 1025x1024", flush=True)
    ds = DeviceScene(renderer, PreparedScene(sc, RenderOptions()))
    # a full-size output, so that nothing could be written out of bounds
    out = torch.zeros(1025 * 1024 * 4, dtype=torch.float32, device="cuda")
    with pytest.raises(RenderError) as e:
        ds.render_noalloc(out)
    assert e.value.kind == "InvalidScene"
    assert "no-allocation" in e.value.message
