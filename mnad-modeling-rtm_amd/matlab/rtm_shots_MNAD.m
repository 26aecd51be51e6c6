function [img_RTM, illum, data, datas] = rtm_shots_MNAD(dx,dz,dt,nt,w,v,vs,xs,zs,xr,zr,d_pml)
% RTM_SHOTS_MNAD  The RTM part of SRC/migration/main_Acoustic_RTM.m:172-262 in
% one batched GPU call (mnad_mex 'rtm_shots'): per shot the observed-data
% forward on v, the forward on the migration model vs keeping its source
% wavefield, the scattered data data - datas (:225) and Born_adj_MNAD with the
% imaging condition; the images and illuminations are stacked on the device
% and img_RTM = mean(Im_s) ./ mean(illum_s) comes back (:243,262).  Inputs as
% the driver's: per-shot cell arrays of 1-based padded grid indices.
% data{ss}, datas{ss} ([nt x nr_ss]) are returned when asked for.
% The source-wavefield storage follows MNAD_SNAP (f16 snapshots by default;
% ring_bf16 / ring_f16 keep a boundary ring and use about 15 times less GPU
% memory per shot).

ns = numel(xs) ;
nrec = zeros(1, ns) ;
for ss = 1:ns
    nrec(ss) = numel(xr{ss}) ;
end
XS = cellfun(@(a) a(1), xs) ;
ZS = cellfun(@(a) a(1), zs) ;
XR = cell2mat(cellfun(@(a) a(:)', xr, 'UniformOutput', false)) ;
ZR = cell2mat(cellfun(@(a) a(:)', zr, 'UniformOutput', false)) ;
if nargout > 2
    [img_RTM, illum, D, Ds] = mnad_mex('rtm_shots', dx, dz, dt, d_pml, v, vs, w, nt, XS, ZS, ...
                                       XR, ZR, nrec, 0) ;
    c = [0, cumsum(nrec)] ;
    data = cell(1, ns) ;
    datas = cell(1, ns) ;
    for ss = 1:ns
        data{ss} = D(:, c(ss)+1:c(ss+1)) ;
        datas{ss} = Ds(:, c(ss)+1:c(ss+1)) ;
    end
else
    [img_RTM, illum] = mnad_mex('rtm_shots', dx, dz, dt, d_pml, v, vs, w, nt, XS, ZS, XR, ZR, ...
                                nrec, 0) ;
end
end
