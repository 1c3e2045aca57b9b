% Linear reaction f(u) = c u: each mode evolves by the closed-form factor
% (1 + dt c) / (1 + dt D l(l+1)) per IMEX Euler step.

function [U, u] = init(noise)
  U = analys(noise);
  u = synth(U);
end

function [Un, u] = step(U, lam, c, D, dt)
  u = synth(U);
  Un = (U + dt * analys(c * u)) ./ (1 + (dt * D) * lam);
end
