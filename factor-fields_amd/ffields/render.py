"""Volume rendering path (filled in below)."""


def attach_render_state(model):
    raise NotImplementedError("reconstruction mode rendering is not built yet")
