% Test model: a linear reaction forced by a fixed grid field (the noise
% buffer) and the analysed state itself.  Its step makes the planner's grid
% op read a host grid buffer and analyse a synthesised value directly.
function [U, u] = init(noise, c)
  U = analys(c + noise);
  u = synth(U);
end

function [Un, u] = step(U, lam, noise, c, D, dt)
  u = synth(U);
  w = u .* noise;
  Un = (U + dt * (analys(w) + c * analys(u))) ./ (1 + (dt * D) * lam);
end
