function data = fwd_shots_MNAD(dx,xs,xr,dz,zs,zr,v,dt,nt,w,d_pml)
% FWD_SHOTS_MNAD  Every shot of the modelling driver's parfor loop
% (SRC/modeling/main_Acoustic_Data.m:169-171) in one batched GPU call
% (mnad_mex 'fwd_shots'): the shots share their launches, the autotuner picks
% the kernel variant and shots per launch.  Same inputs as that loop -- the
% per-shot cell arrays xs{ss}, zs{ss}, xr{ss}, zr{ss} of 1-based padded grid
% indices (main_Acoustic_Data.m:120-126) -- and the same output:
% data{ss} = {datau_ss}, [nt x nr_ss], as fwd_PDE_exact_MNAD returns per shot.
%
%   data = fwd_shots_MNAD(dx,xs,xr,dz,zs,zr,v,dt,nt,w,d_pml)
%   % in place of:  parfor ss = 1:ns, data{ss} = fwd_PDE_exact_MNAD(...); end

ns = numel(xs) ;
nrec = zeros(1, ns) ;
for ss = 1:ns
    nrec(ss) = numel(xr{ss}) ;
end
XS = cellfun(@(a) a(1), xs) ;
ZS = cellfun(@(a) a(1), zs) ;
XR = cell2mat(cellfun(@(a) a(:)', xr, 'UniformOutput', false)) ;
ZR = cell2mat(cellfun(@(a) a(:)', zr, 'UniformOutput', false)) ;
D = mnad_mex('fwd_shots', dx, dz, dt, d_pml, v, w, nt, XS, ZS, XR, ZR, nrec, 0) ;
c = [0, cumsum(nrec)] ;
data = cell(1, ns) ;
for ss = 1:ns
    data{ss} = {D(:, c(ss)+1:c(ss+1))} ;
end
end
