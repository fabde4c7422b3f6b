-- Input deck written for this repository's tests: the 3D shock box of the
-- hydro_3d app (a low-pressure octant x < 0, y < 0, z < 0 inside a gas at
-- rest), in the table layout the app's Lua reader expects.

local states = {
  outside = { rho = 1.0,   p = 1.0 },
  octant  = { rho = 0.125, p = 0.1 },
}

local function region(x, y, z)
  if x < 0 and y < 0 and z < 0 then return "octant" end
  return "outside"
end

hydro = {
  prefix      = "shock_box_3d",
  postfix     = "dat",
  output_freq = 1000000,
  CFL         = 1./3.,
  final_time  = 0.2,
  max_steps   = 12,
  eos = { type = "ideal_gas", gas_constant = 1.4, specific_heat = 1.0 },
  -- density, velocity, pressure at (x, y, z) and time t
  ics = function (x, y, z, t)
    local s = states[region(x, y, z)]
    return s.rho, {0, 0, 0}, s.p
  end,
}
