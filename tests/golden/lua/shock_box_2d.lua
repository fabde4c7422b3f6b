-- Input deck written for this repository's tests: the 2D shock box of the
-- hydro_2d app (a low-pressure quadrant x < 0, y < 0 inside a gas at rest),
-- in the table layout the app's Lua reader expects.

local gas = { gamma = 1.4, cv = 1.0 }

-- (density, pressure) of the two regions
local quiet  = { 1.0,   1.0 }
local pocket = { 0.125, 0.1 }

local function inside(x, y)
  return x < 0 and y < 0
end

-- returns density, velocity, pressure at point (x, y) and time t
local function initial_state(x, y, t)
  local s = quiet
  if inside(x, y) then
    s = pocket
  else
    s = quiet
  end
  return s[1], {0, 0}, s[2]
end

hydro = {
  prefix      = "shock_box_2d",
  postfix     = "dat",
  output_freq = 1e6,
  CFL         = 1./2.,
  final_time  = 0.2,
  max_steps   = 2 * 10,
  eos = {
    type          = "ideal_gas",
    gas_constant  = gas.gamma,
    specific_heat = gas.cv,
  },
  ics = initial_state,
}
