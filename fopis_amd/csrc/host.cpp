// host-side helpers (placeholder)
