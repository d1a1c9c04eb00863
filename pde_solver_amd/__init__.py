"""MI355X (gfx950) HIP solver backend for pde_solver — Python plumbing over the C ABI."""
