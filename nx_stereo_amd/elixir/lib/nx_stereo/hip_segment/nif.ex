defmodule NxStereo.HipSegment.Nif do
  @moduledoc """
  The NIF over libnxs_hip (c_src/nxs_nif.c): `open/1` a context on a HIP device,
  `submit/2` a record without waiting (its reply arrives as
  `{:nxs_reply, ref, binary}`), `close/1` it.
  """
  @on_load :load

  @doc false
  def load do
    path = :filename.join(:code.priv_dir(:nx_stereo), ~c"nxs_nif")
    :erlang.load_nif(path, 0)
  end

  def open(_device), do: :erlang.nif_error(:not_loaded)
  def submit(_context, _record), do: :erlang.nif_error(:not_loaded)
  def close(_context), do: :erlang.nif_error(:not_loaded)
end
