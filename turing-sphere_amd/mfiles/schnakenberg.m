% Schnakenberg activator-substrate system on the unit sphere:
%   u_t = D1 lap(u) + a - u + u^2 v,   v_t = D2 lap(v) + b - u^2 v.
% One IMEX Euler step: reaction explicit on the Gauss grid, diffusion implicit
% and diagonal in spectral space (lam = l(l+1)).  The host provides synth,
% analys, lam, noise and the parameters; spectral fields are real 2 x nlm.

function [U, V, u, v] = init(noise, a, b)
  us = a + b;
  vs = b / (us * us);
  U = analys(us + noise);
  V = analys(vs * ones(numel(noise), 1));
  u = synth(U);
  v = synth(V);
end

function [Un, Vn, u, v] = step(U, V, lam, a, b, D1, D2, dt)
  u = synth(U);
  v = synth(V);
  uuv = u .* u .* v;
  Un = (U + dt * analys(a - u + uuv)) ./ (1 + (dt * D1) * lam);
  Vn = (V + dt * analys(b - uuv))     ./ (1 + (dt * D2) * lam);
end
