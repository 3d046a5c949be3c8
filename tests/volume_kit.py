"""A glass slab for the transmission / volume tests (oracle and GPU)."""
import numpy as np

from rustracer_amd import layouts as L
from rustracer_amd import scenes as S

BASE = (0.9, 0.6, 0.3)


def slab(thickness=1.0, color=(0.8, 0.5, 0.2), dist=2.0, width=48, height=48, spp=16):
    """A glass slab between y = 0 (normal +y) and y = -thickness (normal -y), 100 wide,
    transmission 1, ior 1.5, volume attenuation `color` over `dist`, specular factor 0
    (so no reflection lobe carries weight), seen from y = 5 straight down (no
    anti-aliasing), depth 3, no lights but the 0.01 miss."""
    b = S.Builder()
    m = L.default_material()
    m["base_color"] = (*BASE, 1.0)
    m["metallic_factor"] = 0.0
    m["roughness_factor"] = 0.3
    m["ior"] = 1.5
    m["transmission_exist"] = 1
    m["transmission_factor"] = 1.0
    m["volume_exists"] = 1
    m["attenuation_color"] = color
    m["attenuation_distance"] = dist
    m["specular_factor"] = 0.0
    mid = b.add_material(m)
    geos = []
    for y, up in ((0.0, True), (-thickness, False)):
        s = 50.0
        pos = np.array([[-s, y, -s], [s, y, -s], [s, y, s], [-s, y, s]], np.float64)
        nrm = np.tile([0.0, 1.0 if up else -1.0, 0.0], (4, 1))
        idx = np.array([0, 2, 1, 0, 3, 2] if up else [0, 1, 2, 0, 2, 3], np.uint32)
        geos.append(b.add_primitive(S.make_vertices(pos, nrm), idx, mid))
    inst = S.make_instances([(None, g) for g in geos])
    view = S.look_at_rh((0, 5, 0), (0, 0, 0), (0, 0, -1))
    proj = S.perspective_vk(width / height, 40.0, 0.1, 100.0)
    ubo = S.make_ubo(view, proj, samples=spp, bounces=3, antialiasing=False)
    return b.finish("slab", inst, [], [L.default_dlight()], ubo, width, height)


def blend_layer(alpha=0.3, e_layer=0.4, e_back=0.1, width=48, height=48, spp=16):
    """One alpha-BLEND quad (constant opacity `alpha`, emissive `e_layer`) at y = 0 in
    front of an opaque emissive quad (emissive `e_back`) at y = -2, seen from y = 5
    straight down, depth 1 (only the first hit's emission counts), no lights."""
    b = S.Builder()
    layer = L.default_material()
    layer["alpha_mode"] = L.ALPHA_BLEND
    layer["base_color"] = (0.0, 0.0, 0.0, alpha)
    layer["emissive_factor"][:3] = (e_layer, e_layer, e_layer)
    back = L.default_material()
    back["base_color"] = (0.0, 0.0, 0.0, 1.0)
    back["emissive_factor"][:3] = (e_back, e_back, e_back)
    geos = []
    for y, mid in ((0.0, b.add_material(layer)), (-2.0, b.add_material(back))):
        # the quad's diagonal runs far outside the view, so no camera ray meets two
        # triangles of the layer at their shared edge
        pos = np.array([[-50, y, -50], [50, y, -50], [50, y, 150], [-50, y, 150]], np.float64)
        geos.append(b.add_primitive(S.make_vertices(pos, np.tile([0.0, 1.0, 0.0], (4, 1))),
                                    np.array([0, 2, 1, 0, 3, 2], np.uint32), mid))
    inst = S.make_instances([(None, g) for g in geos])
    view = S.look_at_rh((0, 5, 0), (0, 0, 0), (0, 0, -1))
    proj = S.perspective_vk(width / height, 40.0, 0.1, 100.0)
    ubo = S.make_ubo(view, proj, samples=spp, bounces=1, antialiasing=False, fully_opaque=False)
    return b.finish("blend_layer", inst, [], [L.default_dlight()], ubo, width, height)


# Emission of the near layer, the far layer and the opaque back quad of blend_stack,
# each in a colour channel of its own (red, green, blue).
STACK_EMIT = (0.4, 0.2, 0.1)


def blend_stack(alphas=(0.3, 0.6), width=48, height=48, spp=16):
    """Two alpha-BLEND quads (constant opacities `alphas`) at y = 0 and y = -1 in front
    of an opaque quad at y = -2, seen from y = 5 straight down, depth 1, no lights.
    The near layer emits only red, the far layer only green and the back only blue
    (STACK_EMIT), so each channel of a pixel counts the samples that ended on one
    surface."""
    b = S.Builder()
    geos = []
    for k, y in enumerate((0.0, -1.0, -2.0)):
        m = L.default_material()
        if k < 2:
            m["alpha_mode"] = L.ALPHA_BLEND
            m["base_color"] = (0.0, 0.0, 0.0, alphas[k])
        else:
            m["base_color"] = (0.0, 0.0, 0.0, 1.0)
        e = [0.0, 0.0, 0.0]
        e[k] = STACK_EMIT[k]
        m["emissive_factor"][:3] = e
        mid = b.add_material(m)
        # diagonals far outside the view, as in blend_layer
        pos = np.array([[-50, y, -50], [50, y, -50], [50, y, 150], [-50, y, 150]], np.float64)
        geos.append(b.add_primitive(S.make_vertices(pos, np.tile([0.0, 1.0, 0.0], (4, 1))),
                                    np.array([0, 2, 1, 0, 3, 2], np.uint32), mid))
    inst = S.make_instances([(None, g) for g in geos])
    view = S.look_at_rh((0, 5, 0), (0, 0, 0), (0, 0, -1))
    proj = S.perspective_vk(width / height, 40.0, 0.1, 100.0)
    ubo = S.make_ubo(view, proj, samples=spp, bounces=1, antialiasing=False, fully_opaque=False)
    return b.finish("blend_stack", inst, [], [L.default_dlight()], ubo, width, height)


def stack_counts(img, exposure=5.0):
    """Per-pixel sample counts (near, far, back) of a blend_stack image, as float64
    (the caller checks that they are whole numbers)."""
    e = np.float32(STACK_EMIT) * np.float32(exposure)
    return np.stack([img[..., c] / float(e[c]) for c in range(3)], axis=-1)


def stack_probabilities(alphas):
    """Probability that a sample ends on the near layer, the far layer or the back:
    a1, (1 - a1) a2 and (1 - a1)(1 - a2), whatever order the two any-hit draws are made
    in (a layer is kept only if its own draw keeps it and no nearer surface was kept)."""
    a1, a2 = alphas
    return np.array([a1, (1 - a1) * a2, (1 - a1) * (1 - a2)])
