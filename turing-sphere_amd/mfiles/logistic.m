% Logistic reaction f(u) = r u (1 - u): a uniform field stays uniform and
% follows the scalar explicit-Euler map exactly.

function [U, u] = init(noise)
  U = analys(noise);
  u = synth(U);
end

function [Un, u] = step(U, lam, r, D, dt)
  u = synth(U);
  Un = (U + dt * analys(r * u .* (1 - u))) ./ (1 + (dt * D) * lam);
end
